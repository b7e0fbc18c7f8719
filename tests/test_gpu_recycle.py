"""The recycled start (tos_set_start_recycling) on the GPU against the oracle: the projected
start after a history of solves, and optimisation loops at precision 0 and 3 with the
oracle's compliance per step and fewer iterations than the plain warm start."""
import numpy as np
import pytest

from oracle_api import oracle_api
from test_oracle import E, NU, P, RMIN
from test_oracle_recycle import start_of_next_solve
from topoptspline3d_amd.context import Context
from topoptspline3d_amd.domains import cantilever

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from topoptspline3d_amd.backend import gpu_available, hip_api
    if not gpu_available():
        pytest.skip("no GPU")
    return hip_api()


def make(api, dims, depth):
    solid, support, f = cantilever(*dims)
    c = Context(api, solid, support)
    c.set_material(E, NU, P, RMIN)
    c.set_loads(f)
    c.set_density(np.where(solid.ravel() > 0, 0.4, 0).astype(np.float32))
    c.set_start_recycling(depth)
    return c


def oracle_oc(o):
    """one OC step on the oracle context; returns the new densities"""
    o.sensitivity()
    o.filter(2.0, 0)
    o.oc_update(0.4, 0.05, 0.5)
    return o.density()


def test_projected_start_matches_oracle(hip):
    dims = (64, 32, 32)
    o, g = make(oracle_api(), dims, 4), make(hip, dims, 4)
    for step in range(5):
        for c in (o, g):
            c.update_operator()
            c.solve(rtol=1e-11, maxit=400, precision=0)
        rho = oracle_oc(o)
        g.set_density(rho)
    for c in (o, g):
        c.update_operator()
    uo, ug = start_of_next_solve(o), start_of_next_solve(g)
    d = np.linalg.norm(ug - uo) / np.linalg.norm(uo)
    print(f"\n[recycled start] GPU vs oracle projected start: relative difference {d:.2e}")
    assert d <= 1e-8


@pytest.mark.parametrize("precision", [0, 3])
def test_recycled_loop_matches_oracle(hip, precision):
    dims = (64, 32, 32)
    o = make(oracle_api(), dims, 4)
    gs = {d: make(hip, dims, d) for d in (0, 4)}
    its = {0: [], 4: []}
    for step in range(8):
        o.update_operator()
        o.solve(rtol=1e-8, maxit=400)
        co = o.compliance()[0]
        for d, g in gs.items():
            g.update_operator()
            st = g.solve(rtol=1e-8, maxit=400, precision=precision)
            assert st["rel_res"][0] <= 1e-8
            its[d].append(st["iterations"][0])
            assert abs(g.compliance()[0] - co) <= 1e-7 * co
        rho = oracle_oc(o)
        for g in gs.values():
            g.set_density(rho)
    print(f"\n[recycled loop] precision {precision}: iterations plain {its[0]} recycled {its[4]}")
    assert sum(its[4][3:]) < sum(its[0][3:])
