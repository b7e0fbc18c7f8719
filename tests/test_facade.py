"""The reference-named C++ facade (include/tos_facade.hpp) compiles with g++ and drives
the ordinary SIMP loop; on the CPU it runs over the oracle library, on the GPU over the
HIP library, and both agree with the Python driver."""
import os
import subprocess

import numpy as np
import pytest

from oracle_api import LIB as ORACLE_LIB, oracle_api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "topoptspline3d_amd", "examples", "ordinary_topopt.cpp")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("facade") / "ordinary_topopt")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), SRC, "-o", exe, "-ldl"])
    return exe


def run(exe, lib, prefix, dims, iters):
    out = subprocess.check_output([exe, lib, prefix, *map(str, dims), str(iters)], text=True, timeout=600)
    return np.array([[float(v) for v in line.split()] for line in out.strip().splitlines()])


def test_facade_drives_the_oracle(driver):
    oracle_api()  # builds the library when needed
    rows = run(driver, ORACLE_LIB, "orc_", (24, 12, 12), 4)
    from topoptspline3d_amd.drivers import run_oc
    ref = run_oc(oracle_api(), 24, 12, 12, iters=4, rtol=1e-8)
    assert np.allclose(rows[:, 1], ref["c"], rtol=1e-10)
    assert np.all(rows[:, 4] <= 1e-8)


@pytest.mark.gpu
def test_facade_drives_the_hip_library(driver):
    from topoptspline3d_amd.backend import LIB, gpu_available
    if not gpu_available():
        pytest.skip("no GPU")
    rows = run(driver, LIB, "tos_", (24, 12, 12), 4)
    ref = run(driver, ORACLE_LIB, "orc_", (24, 12, 12), 4)
    assert np.all(np.abs(rows[:, 1] - ref[:, 1]) <= 1e-6 * np.abs(ref[:, 1]))
