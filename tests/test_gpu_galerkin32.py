"""Galerkin precision 1: coarse operators built with FP64 sums into FP32 storage for the
FP32 cycles, FP64 stencils built on demand.  Checks, through the C ABI:
* the on-demand FP64 stencils (and so every FP64 result) are bit-identical to those of
  the default FP64 build, whatever ran before;
* V-cycles on the FP32-built stencils stay within 1e-4 of those on the FP32 copies of
  the FP64 ones, also after a second operator update;
* FP32 / FP16-stencil solves reach the tolerance in the same number of FCG iterations
  (within one) with the same compliance, and an FP64 solve in between changes none of
  their results.
"""
import numpy as np
import pytest

from topoptspline3d_amd.context import Context
from topoptspline3d_amd.domains import cantilever

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from topoptspline3d_amd.backend import gpu_available, hip_api
    if not gpu_available():
        pytest.skip("no GPU")
    return hip_api()


def make(hip, gal, dims=(40, 20, 24), seed=3):
    nx, ny, nz = dims
    solid, support, f = cantilever(nx, ny, nz)
    c = Context(hip, solid, support)
    c.set_material(1e6, 0.4, 3.0, 1e-3)
    c.set_loads(f)
    c.set_galerkin_precision(gal)
    rng = np.random.default_rng(seed)
    c.set_density(rng.uniform(0.01, 1.0, solid.size).astype(np.float32))
    c.update_operator()
    c.solid = solid
    return c


def test_on_demand_fp64_stencils_bit_identical(hip):
    a, b = make(hip, 0), make(hip, 1)
    n = a.info()["n_levels"]
    print(f"\n[gal32] levels {n}", flush=True)
    # FP32 cycle first on b (its FP64 stencils are not built yet), then FP64 stencils
    r = np.random.default_rng(1).normal(size=3 * a.nn)
    za, zb = a.vcycle(r, precision=1), b.vcycle(r, precision=1)
    d = np.abs(zb - za).max() / np.abs(za).max()
    print(f"[gal32] fp32 vcycle rel diff {d:.2e}", flush=True)
    assert d < 1e-4
    for l in range(1, n):
        sa, sb = a.stencil(l), b.stencil(l)
        assert np.array_equal(sa, sb), l
    assert np.array_equal(a.vcycle(r, precision=0), b.vcycle(r, precision=0))


def test_second_update_refreshes_both(hip):
    """A second operator update refreshes the FP32-built stencils (V-cycles of both builds
    stay close) and the on-demand FP64 ones (still bit-identical)."""
    a, b = make(hip, 0, seed=5), make(hip, 1, seed=5)
    r = np.random.default_rng(2).normal(size=3 * a.nn)
    for l in range(1, a.info()["n_levels"]):  # build b's FP64 stencils for the first operator
        assert np.array_equal(a.stencil(l), b.stencil(l)), l
    rho = np.random.default_rng(9).uniform(0.01, 1.0, a.solid.size).astype(np.float32)
    for c in (a, b):
        c.set_density(rho)
        c.update_operator()
    for prec in (1, 2):
        za, zb = a.vcycle(r, precision=prec), b.vcycle(r, precision=prec)
        d = np.abs(zb - za).max() / np.abs(za).max()
        print(f"\n[gal32] precision {prec} vcycle rel diff {d:.2e}", flush=True)
        assert d < 1e-4
    for l in range(1, a.info()["n_levels"]):
        assert np.array_equal(a.stencil(l), b.stencil(l)), l


def test_fp32_built_solves(hip):
    a, b = make(hip, 0, seed=7), make(hip, 1, seed=7)
    out = {}
    for tag, c in (("fp64 build", a), ("fp32 build", b)):
        res = {}
        for prec in (1, 2, 0, 1, 2):
            c.set_displacement(np.zeros(3 * c.nn))
            st = c.solve(rtol=1e-6, maxit=300, precision=prec, warm_start=True)
            assert st["rel_res"][0] <= 1e-6, (tag, prec, st)
            u = c.displacement()
            if prec in res:  # the FP64 solve in between changed nothing
                assert np.array_equal(res[prec][2], u), (tag, prec)
            res[prec] = (st["iterations"][0], c.compliance()[0], u)
        out[tag] = res
        print(f"\n[gal32] {tag}: iterations " + ", ".join(f"p{p} {v[0]}" for p, v in sorted(res.items())),
              flush=True)
    ra, rb = out["fp64 build"], out["fp32 build"]
    assert np.array_equal(ra[0][2], rb[0][2])  # FP64 solves identical
    for prec in (1, 2):
        assert abs(rb[prec][0] - ra[prec][0]) <= 1, prec
        assert abs(rb[prec][1] - ra[prec][1]) <= 1e-6 * abs(ra[prec][1]), prec
