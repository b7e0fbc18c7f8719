"""The skip-layer hierarchy on the GPU against the oracle: the penalised factor-4 Galerkin
level and the dyadic level below it, V-cycles at every precision, solves, and the
reference-shaped 24x12x12 cantilever in its 32^3 cube."""
import numpy as np
import pytest

from oracle_api import oracle_api
from test_oracle import E, NU, P, RMIN
from test_oracle_skip import make
from topoptspline3d_amd.context import Context
from topoptspline3d_amd.domains import cantilever, embed_in_cube

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from topoptspline3d_amd.backend import gpu_available, hip_api
    if not gpu_available():
        pytest.skip("no GPU")
    return hip_api()


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def pair(hip, nx, ny, nz, levels, seed=3, void_frac=0.2, gal=0):
    o, solid, support, f, rho = make(nx, ny, nz, levels, seed=seed, void_frac=void_frac)
    g = Context(hip, solid, support, n_levels=levels, skip_layer=True)
    g.set_material(E, NU, P, RMIN)
    g.set_loads(f)
    g.set_density(rho)
    g.set_galerkin_precision(gal)
    g.update_operator()
    assert g.info()["level_dims"] == o.info()["level_dims"]
    assert g.info()["n_free_nodes"] == o.info()["n_free_nodes"]
    return o, g, f


@pytest.mark.parametrize("dims,levels", [((12, 8, 8), 3), ((64, 32, 32), 0)])
def test_skip_stencils_match_oracle(hip, dims, levels):
    o, g, _ = pair(hip, *dims, levels)
    for lev in range(1, o.info()["n_levels"]):
        so, sg = o.stencil(lev), g.stencil(lev)
        print(f"\n[skip stencil] {dims} level {lev}: relative difference {rel(sg, so):.2e}")
        assert rel(sg, so) <= 1e-12


def test_skip_vcycle_matches_oracle(hip):
    o, g, f = pair(hip, 64, 32, 32, 0, seed=5, void_frac=0.15, gal=1)
    assert o.info()["n_levels"] >= 3
    fr = np.tile((o.node_flags(0) & 4) > 0, 3)
    r = np.where(fr, np.random.default_rng(2).normal(size=3 * o.nn), 0)
    zo = o.vcycle(r)
    for prec, tol in ((0, 1e-9), (1, 2e-4), (2, 1e-2)):
        zg = g.vcycle(r, precision=prec)
        print(f"[skip vcycle] precision {prec}: relative difference {rel(zg, zo):.2e}")
        assert rel(zg, zo) <= tol


@pytest.mark.parametrize("precision", [0, 2, 3])
def test_skip_solve_matches_oracle(hip, precision):
    o, g, f = pair(hip, 64, 32, 32, 0, seed=7, gal=0 if precision == 0 else 1)
    so = o.solve(rtol=1e-8, maxit=400)
    sg = g.solve(rtol=1e-8, maxit=400, precision=precision)
    print(f"\n[skip solve] precision {precision}: GPU {sg['iterations'][0]} iterations, oracle {so['iterations'][0]}")
    assert sg["rel_res"][0] <= 1e-8
    co, cg = o.compliance()[0], g.compliance()[0]
    assert abs(cg - co) <= 1e-7 * co
    if precision == 0:
        assert abs(sg["iterations"][0] - so["iterations"][0]) <= 1


def test_reference_shaped_c0_trajectory(hip):
    """ten OC iterations of the 24x12x12 cantilever in the reference's 32^3 cube with the
    skip layer (the reference's own hierarchy shape): compliance per iteration within 1e-6
    of the oracle's"""
    solid, support, f = cantilever(24, 12, 12)
    S, U, F, _, off = embed_in_cube(solid, support, f)
    cs = {}
    for tag, api in (("oracle", oracle_api()), ("gpu", hip)):
        c = Context(api, S, U, skip_layer=True)
        c.set_material(E, NU, P, RMIN)
        c.set_loads(F)
        c.set_density(np.where(S.ravel() > 0, 0.4, 0).astype(np.float32))
        assert c.info()["n_free_nodes"][1] == 175
        out = []
        for it in range(10):
            c.update_operator()
            c.solve(rtol=1e-8, maxit=300, precision=0)
            out.append(c.compliance()[0])
            c.sensitivity()
            c.filter(2.0, 0)
            c.oc_update(0.4, 0.03, 0.5)
        cs[tag] = np.array(out)
    print(f"\n[reference-shaped C0] compliance {cs['gpu'][[0, -1]]} (oracle {cs['oracle'][[0, -1]]})")
    assert np.all(np.abs(cs["gpu"] - cs["oracle"]) <= 1e-6 * cs["oracle"])
