"""The oracle's recycled start (tos_set_start_recycling): with one basis vector it is the
energy-optimal scaling of the warm start (closed form); with more it satisfies the
Galerkin condition V^T (f - K u0) = 0 on span{u, s_2, ..., s_m} and lowers the energy
functional below the warm start's; in an optimisation loop it keeps the trajectory and
needs fewer iterations."""
import numpy as np

from oracle_api import oracle_api
from test_oracle import E, NU, P, RMIN, free_mask
from topoptspline3d_amd.context import Context
from topoptspline3d_amd.domains import cantilever


def loop_ctx(api, nx=16, ny=8, nz=8, depth=0):
    solid, support, f = cantilever(nx, ny, nz)
    c = Context(api, solid, support)
    c.set_material(E, NU, P, RMIN)
    c.set_loads(f)
    c.set_density(np.where(solid.ravel() > 0, 0.4, 0).astype(np.float32))
    c.set_start_recycling(depth)
    return c


def oc_step(c):
    c.sensitivity()
    c.filter(1.5, 0)
    c.oc_update(0.4, 0.05, 0.5)
    c.update_operator()


def energy(c, u):
    f = c.get_loads(0)
    return 0.5 * u @ c.apply(u) - f @ u


def start_of_next_solve(c):
    """the projected start: a solve with no iteration leaves it as the displacement"""
    st = c.solve(rtol=1e-30, maxit=0, check=False)
    assert st["rc"] != 0  # not converged, so the history is not extended
    return c.displacement()


def test_depth_one_is_the_optimal_scaling_of_the_warm_start():
    c = loop_ctx(oracle_api(), depth=1)
    c.update_operator()
    fm = np.tile(free_mask(c), 3)
    u = np.where(fm, np.random.default_rng(1).normal(size=fm.size), 0.0)
    c.set_displacement(u)
    f = c.get_loads(0)
    u0 = start_of_next_solve(c)
    scale = (u @ f) / (u @ c.apply(u))
    assert np.allclose(u0, scale * u, rtol=1e-12, atol=1e-14 * np.abs(u0).max())


def test_galerkin_condition_on_the_stored_solutions():
    c = loop_ctx(oracle_api(), depth=4)
    c.update_operator()
    sols = []
    for _ in range(5):
        c.solve(rtol=1e-10, maxit=300)
        sols.append(c.displacement())
        oc_step(c)
    # basis: the warm start (= newest solution) and the three before it
    V = np.stack([sols[-1], sols[-2], sols[-3], sols[-4]], 1)
    u_cur = sols[-1]
    u0 = start_of_next_solve(c)
    f = c.get_loads(0)
    coef, *_ = np.linalg.lstsq(V, u0, rcond=None)
    assert np.linalg.norm(V @ coef - u0) <= 1e-10 * np.linalg.norm(u0)  # u0 lies in the span
    r0 = f - c.apply(u0)
    g = V.T @ r0
    assert np.abs(g).max() <= 1e-9 * np.linalg.norm(V, axis=0).max() * np.linalg.norm(f)
    assert energy(c, u0) < energy(c, u_cur)
    # the unconverged call leaves the history as it was; u0 lies in the same span, so the
    # next projection lands on u0 again (to the rounding of the Gram solve)
    assert np.linalg.norm(start_of_next_solve(c) - u0) <= 1e-8 * np.linalg.norm(u0)


def test_recycled_loop_keeps_the_trajectory_with_fewer_iterations():
    out = {}
    for depth in (0, 4):
        c = loop_ctx(oracle_api(), 24, 12, 12, depth=depth)
        c.update_operator()
        its, comp = [], []
        for _ in range(10):
            st = c.solve(rtol=1e-8, maxit=300)
            its.append(st["iterations"][0])
            comp.append(c.compliance()[0])
            oc_step(c)
        out[depth] = (np.array(its), np.array(comp))
    print(f"\n[recycled start] iterations plain {out[0][0].tolist()} recycled {out[4][0].tolist()}")
    assert np.all(np.abs(out[4][1] - out[0][1]) <= 1e-7 * out[0][1])
    assert out[4][0][3:].sum() < out[0][0][3:].sum()
    assert np.all(out[4][0] <= out[0][0] + 1)
