"""Records of the reference's default loop written from the GPU library (utils/binaryIO.h
layout, optimization/optimization.cpp:728-765): the reference's own reader (read_data/
main.cpp, built from the reference tree into oracle/_ref) prints the GPU run's records,
the records hold the GPU context's vectors, and they match the oracle run's records."""
import os
import re
import subprocess

import numpy as np
import pytest

from oracle_api import ROOT, oracle_api
from topoptspline3d_amd import bio
from topoptspline3d_amd.drivers import run_worst_oc

pytestmark = pytest.mark.gpu
READ_DATA = os.path.join(ROOT, "oracle", "_ref", "read_data")


@pytest.fixture(scope="module")
def hip():
    from topoptspline3d_amd.backend import gpu_available, hip_api
    if not gpu_available():
        pytest.skip("no GPU")
    return hip_api()


def test_gpu_records(hip, tmp_path):
    g, o = tmp_path / "gpu", tmp_path / "oracle"
    g.mkdir(); o.mkdir()
    rg = run_worst_oc(hip, 24, 12, 12, iters=3, out_dir=str(g))
    ro = run_worst_oc(oracle_api(), 24, 12, 12, iters=3, out_dir=str(o))
    ctx, ids = rg["ctx"], rg["load_nodes"]
    f = bio.read_vector(str(g / "flast")).reshape(-1, 3)
    assert np.array_equal(f, ctx.get_loads(0).reshape(3, -1)[:, ids].T)
    u = bio.read_vector(str(g / "ulast")).reshape(-1, 3)
    assert np.array_equal(u, ctx.displacement(0).reshape(3, -1).T)
    assert np.array_equal(bio.read_vector(str(g / "cworst")), rg["c"])
    for name in ("cworst", "vrec"):
        a, b = bio.read_vector(str(g / name)), bio.read_vector(str(o / name))
        assert np.allclose(a, b, rtol=1e-6, atol=0), name
    fo = bio.read_vector(str(o / "flast"))
    assert min(np.linalg.norm(f.ravel() - fo), np.linalg.norm(f.ravel() + fo)) <= 1e-4 * np.linalg.norm(fo)
    if not os.path.exists(READ_DATA):
        pytest.skip("oracle/_ref/read_data not built (reference tree absent where the suite was built)")
    for name, key in (("vrec", "v"), ("cworst", "c")):
        out = subprocess.check_output([READ_DATA, str(g / name)], text=True, timeout=60)
        lines = out.strip().splitlines()
        assert int(lines[0]) == 3
        vals = [float(re.search(r"v =\s*(\S+)", l).group(1)) for l in lines[1:]]
        assert np.allclose(vals, rg[key], rtol=0, atol=5e-7), (name, vals, rg[key])
