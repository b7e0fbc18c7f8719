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
SRC_WORST = os.path.join(ROOT, "topoptspline3d_amd", "examples", "worst_case_topopt.cpp")


def _compile(tmp_path_factory, src, name):
    exe = str(tmp_path_factory.mktemp("facade") / name)
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), src, "-o", exe, "-ldl"])
    return exe


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    return _compile(tmp_path_factory, SRC, "ordinary_topopt")


@pytest.fixture(scope="module")
def worst_driver(tmp_path_factory):
    return _compile(tmp_path_factory, SRC_WORST, "worst_case_topopt")


def run(exe, lib, prefix, dims, iters, *extra):
    out = subprocess.check_output([exe, lib, prefix, *map(str, dims), str(iters), *map(str, extra)], text=True,
                                  timeout=600)
    return np.array([[float(v) for v in line.split()] for line in out.strip().splitlines()])


def test_facade_drives_the_oracle(driver):
    oracle_api()  # builds the library when needed
    rows = run(driver, ORACLE_LIB, "orc_", (24, 12, 12), 4)
    from topoptspline3d_amd.drivers import run_oc
    ref = run_oc(oracle_api(), 24, 12, 12, iters=4, rtol=1e-8)
    assert np.allclose(rows[:, 1], ref["c"], rtol=1e-10)
    assert np.all(rows[:, 4] <= 1e-8)


def test_facade_recycled_start_on_the_oracle(driver):
    """getGrids().set_start_recycling(4) through the facade: the same compliance trajectory
    with no more iterations than the plain warm start"""
    oracle_api()
    plain = run(driver, ORACLE_LIB, "orc_", (24, 12, 12), 8)
    rec = run(driver, ORACLE_LIB, "orc_", (24, 12, 12), 8, 4)
    assert np.all(np.abs(rec[:, 1] - plain[:, 1]) <= 1e-7 * plain[:, 1])
    assert rec[:, 3].sum() < plain[:, 3].sum() and np.all(rec[:, 4] <= 1e-8)


@pytest.mark.gpu
def test_facade_drives_the_hip_library(driver):
    from topoptspline3d_amd.backend import LIB, gpu_available
    if not gpu_available():
        pytest.skip("no GPU")
    rows = run(driver, LIB, "tos_", (24, 12, 12), 4)
    ref = run(driver, ORACLE_LIB, "orc_", (24, 12, 12), 4)
    assert np.all(np.abs(rows[:, 1] - ref[:, 1]) <= 1e-6 * np.abs(ref[:, 1]))


def test_facade_worst_case_loop_on_the_oracle(worst_driver):
    """optimization()'s loop (modifiedPM + sensitivity + OC) through the facade equals the
    Python driver over the same library (to the 13 printed digits)."""
    oracle_api()
    rows = run(worst_driver, ORACLE_LIB, "orc_", (16, 8, 8), 3)
    from topoptspline3d_amd.drivers import run_worst_oc
    # the facade's Parameter holds float fields, as the reference's (optimization.h:15-35)
    ref = run_worst_oc(oracle_api(), 16, 8, 8, iters=3, nu=float(np.float32(0.4)),
                       rho_min=float(np.float32(1e-3)))
    assert np.allclose(rows[:, 1], ref["c"], rtol=1e-11, atol=0)  # printed with 13 digits
    assert np.array_equal(rows[:, 3].astype(int), ref["its"])
    assert np.all(rows[:, 3] < 500)


@pytest.mark.gpu
def test_facade_worst_case_loop_on_the_hip_library(worst_driver):
    from topoptspline3d_amd.backend import LIB, gpu_available
    if not gpu_available():
        pytest.skip("no GPU")
    rows = run(worst_driver, LIB, "tos_", (16, 8, 8), 3)
    ref = run(worst_driver, ORACLE_LIB, "orc_", (16, 8, 8), 3)
    assert np.all(np.abs(rows[:, 1] - ref[:, 1]) <= 1e-6 * np.abs(ref[:, 1]))
    assert np.all(np.abs(rows[:, 3] - ref[:, 3]) <= 1)


SRC_SS = os.path.join(ROOT, "topoptspline3d_amd", "examples", "self_supporting_topopt.cpp")
SS_DIMS, SS_ITERS = (16, 8, 8), 4


@pytest.fixture(scope="module")
def ss_driver(tmp_path_factory):
    return _compile(tmp_path_factory, SRC_SS, "self_supporting_topopt")


def _ss_coeffs(tmp_path_factory):
    """wavy initial coefficients (so the spline has an interior surface from the start)"""
    from topoptspline3d_amd.context import Context
    from topoptspline3d_amd.domains import cantilever
    solid, support, _ = cantilever(*SS_DIMS)
    ctx = Context(oracle_api(), solid, support)
    n = ctx.spline_init(1, 4, 4)
    i = np.arange(n)
    c0 = (0.4 + 0.3 * np.sin(0.9 * i) * np.cos(0.37 * i)).astype(np.float32)
    path = str(tmp_path_factory.mktemp("coeffs") / "c0.f32")
    c0.tofile(path)
    return c0, path


def _run_ss(exe, lib, prefix, path):
    out = subprocess.check_output([exe, lib, prefix, *map(str, SS_DIMS), str(SS_ITERS), path], text=True,
                                  timeout=600)
    return np.array([[float(v) for v in line.split()] for line in out.strip().splitlines()])


def test_facade_self_supporting_loop_on_the_oracle(ss_driver, tmp_path_factory):
    """the self-supporting spline loop (Heaviside projection, surface constraints, MMA
    m = 3) through the facade equals the Python driver over the same library"""
    c0, path = _ss_coeffs(tmp_path_factory)
    rows = _run_ss(ss_driver, ORACLE_LIB, "orc_", path)
    from topoptspline3d_amd.drivers import run_spline_mma
    f32 = lambda v: float(np.float32(v))  # the facade's Parameter holds floats (optimization.h:15-35)
    ref = run_spline_mma(oracle_api(), *SS_DIMS, iters=SS_ITERS, rtol=1e-8, precision=0, heaviside=True,
                         self_support=True, c0=c0, vfrac=f32(0.4), nu=f32(0.4), rho_min=f32(1e-3))
    assert rows.shape == (SS_ITERS, 7)
    assert np.allclose(rows[:, 1], ref["c"], rtol=1e-11, atol=0)
    assert np.allclose(rows[:, 3], ref["ss"], rtol=1e-11, atol=1e-14)
    assert np.allclose(rows[:, 4], ref["drip"], rtol=1e-11, atol=1e-14)
    assert np.all(rows[:, 3] != 1.0)  # the surface exists from the first iteration


@pytest.mark.gpu
def test_facade_self_supporting_loop_on_the_hip_library(ss_driver, tmp_path_factory):
    from topoptspline3d_amd.backend import LIB, gpu_available
    if not gpu_available():
        pytest.skip("no GPU")
    _, path = _ss_coeffs(tmp_path_factory)
    rows = _run_ss(ss_driver, LIB, "tos_", path)
    ref = _run_ss(ss_driver, ORACLE_LIB, "orc_", path)
    assert np.all(np.abs(rows[:, 1] - ref[:, 1]) <= 1e-5 * np.abs(ref[:, 1]))
    assert np.all(np.abs(rows[:, 3] - ref[:, 3]) <= 1e-4 * np.abs(ref[:, 3]) + 1e-6)
    assert np.all(np.abs(rows[:, 4] - ref[:, 4]) <= 1e-4 * np.abs(ref[:, 4]) + 1e-6)
