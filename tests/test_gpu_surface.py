"""Self-supporting and drip constraints on spline surface points (tos_surface_constraints)
on the GPU against the FP64 oracle, through the C ABI: the same surface points (count and
positions), normals, orientation, global values, and coefficient gradients, for every
self-support mode, the drip modes, both spline layouts (uniform lattice and the
reference's clamped knots); and the three-constraint MMA driver for a few iterations."""
import numpy as np
import pytest

from oracle_api import oracle_api
from topoptspline3d_amd.context import Context
from topoptspline3d_amd.domains import cantilever

pytestmark = pytest.mark.gpu

E, NU, P, RMIN = 1e6, 0.4, 3.0, 1e-3


@pytest.fixture(scope="module")
def hip():
    from topoptspline3d_amd.backend import gpu_available, hip_api
    if not gpu_available():
        pytest.skip("no GPU")
    return hip_api()


def ball(i, j, k, ci, cj, ck, r2):
    return (0.95 - 0.45 / r2 * ((i - ci) ** 2 + (j - cj) ** 2 + (k - ck) ** 2)
            + 0.06 * np.sin(0.9 * i + 0.3) * np.cos(0.7 * j - 0.2) + 0.04 * np.cos(0.8 * k + 0.1 * i))


def pair(hip, nx, ny, nz, mode, order, param):
    solid, support, f = cantilever(nx, ny, nz)
    out = []
    for api in (oracle_api(), hip):
        c = Context(api, solid, support)
        c.set_material(E, NU, P, RMIN)
        n = c.spline_init(mode, order, param)
        out.append(c)
    return out, n


def coeffs_for(ctx_o, nx, ny, nz, mode, order, param):
    n = ctx_o.n_coeffs
    if mode == 1:
        ncx, ncy, ncz = (nx + param - 1) // param + 1, (ny + param - 1) // param + 1, (nz + param - 1) // param + 1
    else:
        ncx = ncy = ncz = order + param
    assert ncx * ncy * ncz == n
    k, j, i = np.meshgrid(np.arange(ncz), np.arange(ncy), np.arange(ncx), indexing="ij")
    r2 = (0.35 * min(ncx, ncy, ncz)) ** 2
    return ball(i, j, k, (ncx - 1) / 2, (ncy - 1) / 2, (ncz - 1) / 2, r2).astype(np.float32).ravel()


CASES = [(1, 4, 4, dict(ss_mode=m, drip_mode=4)) for m in range(6)] + [
    (1, 4, 4, dict(ss_mode=4, drip_mode=6, drip_beta=4.0, drip_mu=2.0, drip_beta_scale=1.0)),
    (1, 4, 4, dict(ss_mode=5, drip_mode=7, drip_beta=4.0, drip_version=1, drip_alpha=0.5, drip_beta_scale=1.0)),
    (1, 4, 4, dict(ss_mode=4, drip_mode=7)),  # the reference's drip parameters
    (0, 3, 6, dict(ss_mode=4, drip_mode=6, drip_beta=4.0, drip_mu=2.0, drip_beta_scale=1.0)),
    (0, 4, 5, dict(ss_mode=2, drip_mode=4)),
]


@pytest.mark.parametrize("mode,order,param,opts", CASES)
def test_surface_constraints_match_oracle(hip, mode, order, param, opts):
    nx, ny, nz = 28, 20, 24
    (o, g), n = pair(hip, nx, ny, nz, mode, order, param)
    cf = coeffs_for(o, nx, ny, nz, mode, order, param)
    kw = dict(cells=20, beta=8.0)
    kw.update(opts)
    out = []
    for c in (o, g):
        c.set_coeffs(cf)
        out.append(c.surface_constraints(**kw))
    so, sg = out
    print(so, sg)
    assert so["n_points"] == sg["n_points"] > 100 and so["n_down"] == sg["n_down"]
    assert so["direction"] == sg["direction"]
    po, no_ = o.surface_points(so["n_points"])
    pg, ng = g.surface_points(sg["n_points"])
    assert np.abs(pg - po).max() <= 1e-9
    assert np.abs(ng - no_).max() <= 1e-9 * np.abs(no_).max()
    for key in ("ss_value", "drip_value"):
        assert abs(sg[key] - so[key]) <= 1e-9 * max(1.0, abs(so[key])), key
    for which in (2, 3):
        go, gg = o.coeff_sens(which), g.coeff_sens(which)
        scale = max(np.abs(go).max(), 1e-30)
        assert np.abs(gg - go).max() <= 1e-5 * scale, which


@pytest.mark.parametrize("cells", [64, 150])
def test_surface_constraints_at_finer_lattices(hip, cells):
    """the reference's 64-cell background lattice and a finer one (the lattice counts are
    scanned over 140 and 1700 blocks of 2048)"""
    nx, ny, nz = 28, 20, 24
    (o, g), n = pair(hip, nx, ny, nz, 1, 4, 4)
    cf = coeffs_for(o, nx, ny, nz, 1, 4, 4)
    out = []
    for c in (o, g):
        c.set_coeffs(cf)
        out.append(c.surface_constraints(cells=cells, beta=8.0))
    so, sg = out
    assert so["n_points"] == sg["n_points"] > 1000
    po, _ = o.surface_points(so["n_points"])
    pg, _ = g.surface_points(sg["n_points"])
    assert np.abs(pg - po).max() <= 1e-9
    assert abs(sg["ss_value"] - so["ss_value"]) <= 1e-9 * max(1.0, abs(so["ss_value"]))
    go, gg = o.coeff_sens(2), g.coeff_sens(2)
    assert np.abs(gg - go).max() <= 1e-5 * max(np.abs(go).max(), 1e-30)


def test_equal_coefficients_give_the_drivers_constants(hip):
    (o, g), n = pair(hip, 16, 12, 12, 1, 4, 4)
    for c in (o, g):
        c.set_coeffs(np.full(n, 0.4, np.float32))
        st = c.surface_constraints(cells=16)
        assert st["n_points"] == 0 and st["ss_value"] == 1.0 and st["drip_value"] == 0.0
        assert not np.any(c.coeff_sens(2)) and not np.any(c.coeff_sens(3))


def test_self_supporting_mma_driver_matches_oracle(hip):
    """utils/test_utils.cpp:972-1217 with ENABLE_SELFSUPPORT on a 32x16x16 cantilever:
    5 iterations, compliance and constraint values within 1e-5 of the oracle's."""
    from topoptspline3d_amd.drivers import run_spline_mma
    kw = dict(iters=5, rtol=1e-8, self_support=True, surface=dict(cells=16))
    ro = run_spline_mma(oracle_api(), 32, 16, 16, precision=0, **kw)
    rg = run_spline_mma(hip, 32, 16, 16, precision=0, **kw)
    print("c", rg["c"], ro["c"], "ss", rg["ss"], ro["ss"])
    assert np.all(np.abs(np.array(rg["c"]) - np.array(ro["c"])) <= 1e-5 * np.abs(np.array(ro["c"])))
    assert np.allclose(rg["ss"], ro["ss"], rtol=1e-5, atol=1e-7)
    assert np.allclose(rg["drip"], ro["drip"], rtol=1e-5, atol=1e-7)
    assert max(ro["ss"]) > 1.0  # surfaces appeared and were constrained
