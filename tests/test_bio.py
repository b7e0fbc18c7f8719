"""Binary records in the reference's `bio` format (utils/binaryIO.h:70-111):
* byte layout of write_vector / write_vectors against struct.pack of known values;
* read_vector reads floor(size / item) items, as the reference's reader;
* the reference's own record reader (read_data/main.cpp, built from the reference tree
  into oracle/_ref by `make -C oracle ref`) prints the records the driver writes;
* the C++ facade (include/tos_bio.hpp, HierarchyGrid::write*) and the Python driver
  write the same bytes for the same worst-case loop over the oracle."""
import os
import re
import struct
import subprocess

import numpy as np
import pytest

from oracle_api import LIB as ORACLE_LIB, ROOT, oracle_api
from topoptspline3d_amd import bio
from topoptspline3d_amd.drivers import run_worst_oc

READ_DATA = os.path.join(ROOT, "oracle", "_ref", "read_data")
NU32, RMIN32 = float(np.float32(0.4)), float(np.float32(1e-3))  # the facade's float parameters


def test_byte_layout(tmp_path):
    p = str(tmp_path / "v")
    bio.write_vector(p, [1.0, -2.5, 3e-300])
    assert open(p, "rb").read() == struct.pack("<3d", 1.0, -2.5, 3e-300)
    bio.write_vector(p, [0.25, -1.0], np.float32)
    assert open(p, "rb").read() == struct.pack("<2f", 0.25, -1.0)
    bio.write_vectors(p, [[1, 2], [3, 4], [5, 6]])
    assert open(p, "rb").read() == struct.pack("<6d", 1, 3, 5, 2, 4, 6)
    with pytest.raises(ValueError):
        bio.write_vectors(p, [[1, 2], [3]])


def test_read_vector_floor_and_round_trip(tmp_path):
    p = str(tmp_path / "v")
    with open(p, "wb") as fh:
        fh.write(struct.pack("<2d", 1.5, 2.5) + b"\x01\x02\x03")  # a partial trailing item is dropped
    assert np.array_equal(bio.read_vector(p), [1.5, 2.5])
    x = np.random.default_rng(0).normal(size=17).astype(np.float32)
    bio.write_vector(p, x)
    assert np.array_equal(bio.read_vector(p, np.float32), x)


def test_reference_reader_prints_driver_records(tmp_path):
    if not os.path.exists(READ_DATA):
        pytest.skip("oracle/_ref/read_data not built (reference tree absent)")
    rec = run_worst_oc(oracle_api(), 12, 6, 6, iters=3, out_dir=str(tmp_path))
    for name, key in (("vrec", "v"), ("cworst", "c")):
        out = subprocess.check_output([READ_DATA, str(tmp_path / name)], text=True, timeout=60)
        lines = out.strip().splitlines()
        assert int(lines[0]) == 3
        vals = [float(re.search(r"v =\s*(\S+)", l).group(1)) for l in lines[1:]]
        assert np.allclose(vals, rec[key], rtol=0, atol=5e-7), (name, vals, rec[key])
    for i in range(1, 4):
        assert os.path.getsize(tmp_path / f"iter{i}_fs") == 3 * 8 * rec["load_nodes"].size


def test_records_hold_the_context_vectors(tmp_path):
    rec = run_worst_oc(oracle_api(), 12, 6, 6, iters=2, out_dir=str(tmp_path))
    ctx, ids = rec["ctx"], rec["load_nodes"]
    f = bio.read_vector(str(tmp_path / "flast")).reshape(-1, 3)
    assert np.array_equal(f, ctx.get_loads(0).reshape(3, -1)[:, ids].T)
    u = bio.read_vector(str(tmp_path / "ulast")).reshape(-1, 3)
    assert np.array_equal(u, ctx.displacement(0).reshape(3, -1).T)
    assert np.array_equal(bio.read_vector(str(tmp_path / "cworst")), rec["c"])
    assert np.array_equal(bio.read_vector(str(tmp_path / "crec_iter")), rec["c"])
    assert bio.read_vector(str(tmp_path / "trec")).size == 2


def test_facade_writes_the_same_records(tmp_path):
    exe = str(tmp_path / "worst_case_topopt")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "topoptspline3d_amd", "examples", "worst_case_topopt.cpp"),
                           "-o", exe, "-ldl"])
    oracle_api()
    a, b = tmp_path / "facade", tmp_path / "python"
    a.mkdir(); b.mkdir()
    subprocess.check_call([exe, ORACLE_LIB, "orc_", "12", "6", "6", "2", str(a)], timeout=600,
                          stdout=subprocess.DEVNULL)
    run_worst_oc(oracle_api(), 12, 6, 6, iters=2, out_dir=str(b), nu=NU32, rho_min=RMIN32)
    for name in ("cworst", "vrec", "flast", "ulast"):
        assert open(a / name, "rb").read() == open(b / name, "rb").read(), name
