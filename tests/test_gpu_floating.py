"""Floating bodies on the GPU against the oracle: forceProject (tos_balance_loads), the
solve with the rigid-motion-regularised coarsest factor, displacementProject
(tos_remove_rigid) and the worst-case search without supports, free and along normals."""
import numpy as np
import pytest

from oracle_api import oracle_api
from test_oracle import E, NU, P, RMIN
from test_oracle_floating import coords, floating, region_force, resultant, rigid_matrix, set_region_force
from topoptspline3d_amd.context import Context, TosError
from topoptspline3d_amd.domains import cantilever, face_load_region

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from topoptspline3d_amd.backend import gpu_available, hip_api
    if not gpu_available():
        pytest.skip("no GPU")
    return hip_api()


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def pair(hip, nx, ny, nz, seed, levels=0):
    o, solid, rho = floating(nx, ny, nz, seed=seed, levels=levels)
    g = Context(hip, solid, None, n_levels=levels)
    g.set_material(E, NU, P, RMIN)
    g.set_density(rho)
    g.set_loads(np.zeros(3 * g.nn))
    g.update_operator()
    assert g.info()["has_support"] is False and g.info()["n_levels"] == o.info()["n_levels"]
    return o, g


@pytest.mark.parametrize("with_normals", [False, True])
def test_balance_loads_matches_oracle(hip, with_normals):
    o, g = pair(hip, 16, 8, 8, seed=1, levels=2)
    ids, nrm = face_load_region(16, 8, 8, "x+")
    ids2, nrm2 = face_load_region(16, 8, 8, "z-", inset=1)
    ids, nrm = np.r_[ids, ids2], np.r_[nrm, nrm2]
    fs = np.random.default_rng(3).normal(size=3 * ids.size)
    for c in (o, g):
        set_region_force(c, ids, fs)
        c.balance_loads(ids, normals=nrm if with_normals else None)
    fo, fg = o.get_loads(0), g.get_loads(0)
    assert np.allclose(fg, fo, rtol=0, atol=1e-12 * np.abs(fs).max())
    F, M = resultant(coords(g, ids), region_force(g, ids))
    assert np.abs(F).max() <= 1e-11 and np.abs(M).max() <= 1e-9


def test_balance_loads_with_supports_matches_oracle(hip):
    solid, support, f = cantilever(16, 8, 8)
    ids, nrm = face_load_region(16, 8, 8, "x+")
    full = np.random.default_rng(4).normal(size=(3, (17 * 9 * 9)))
    full[:, support.ravel() > 0] = 0
    out = []
    for api in (oracle_api(), hip):
        c = Context(api, solid, support)
        c.set_material(E, NU, P, RMIN)
        c.set_loads(full.ravel())
        c.balance_loads(ids, normals=nrm)
        out.append(c.get_loads(0))
        with pytest.raises(TosError):
            c.remove_rigid()
    assert np.allclose(out[1], out[0], rtol=0, atol=1e-14)


@pytest.mark.parametrize("precision", [0, 2, 3])
def test_floating_solve_matches_oracle(hip, precision):
    """a free 64x32x32 block (4 levels) under a balanced random load on two faces"""
    nx, ny, nz = 64, 32, 32
    o, g = pair(hip, nx, ny, nz, seed=2)
    ids, _ = face_load_region(nx, ny, nz, "x+")
    ids0, _ = face_load_region(nx, ny, nz, "x-")
    both = np.r_[ids, ids0]
    rng = np.random.default_rng(6)
    f = np.zeros((3, o.nn))
    f[:, both] = rng.normal(size=(3, both.size))
    for c in (o, g):
        c.set_loads(f.ravel())
        c.balance_loads(both)
    assert np.allclose(g.get_loads(0), o.get_loads(0), rtol=0, atol=1e-12)
    so = o.solve(rtol=1e-9, maxit=300)
    if precision:
        g.set_galerkin_precision(1)
        g.update_operator()
    sg = g.solve(rtol=1e-9, maxit=300, precision=precision)
    print(f"\n[floating solve] precision {precision}: GPU {sg['iterations'][0]} iterations "
          f"(oracle {so['iterations'][0]}), rel_res {sg['rel_res'][0]:.2e}")
    assert sg["rel_res"][0] <= 1e-9
    co, cg = o.compliance()[0], g.compliance()[0]
    assert abs(cg - co) <= 1e-8 * abs(co)
    for c in (o, g):
        c.remove_rigid()
    assert rel(g.displacement(), o.displacement()) <= 1e-7
    ur = g.displacement().reshape(3, -1).T.ravel()
    assert np.abs(rigid_matrix(coords(g, np.arange(g.nn))).T @ ur).max() <= 1e-9 * np.abs(ur).max() * g.nn


@pytest.mark.parametrize("with_normals", [False, True])
def test_floating_worst_case_matches_oracle(hip, with_normals):
    nx, ny, nz = 32, 16, 16
    o, g = pair(hip, nx, ny, nz, seed=7)
    ids, nrm = face_load_region(nx, ny, nz, "x+")
    ids2, nrm2 = face_load_region(nx, ny, nz, "x-")
    ids, nrm = np.r_[ids, ids2], np.r_[nrm, nrm2]
    kw = dict(normals=nrm if with_normals else None, seed=5, max_iter=3000, fch_tol=1e-9, res_tol=1e-11)
    so = o.worst_compliance(ids, **kw)
    sg = g.worst_compliance(ids, precision=0, **kw)
    print(f"\n[floating worst] normals={with_normals}: oracle {so['compliance']:.10e} ({so['iterations']} "
          f"cycles), GPU {sg['compliance']:.10e} ({sg['iterations']} cycles)")
    assert abs(sg["compliance"] - so["compliance"]) <= 1e-8 * so["compliance"]
    fo, fg = region_force(o, ids), region_force(g, ids)
    assert min(rel(fg, fo), rel(-fg, fo)) <= 1e-5
    F, M = resultant(coords(g, ids), fg)
    assert np.abs(F).max() <= 1e-11 and np.abs(M).max() <= 1e-9
    # the reference's stop rule with the mixed-precision cycles lands on the same load
    sm = g.worst_compliance(ids, normals=kw["normals"], seed=5, precision=2)
    assert sm["fch"] <= 1e-4 and abs(sm["compliance"] - so["compliance"]) <= 1e-4 * so["compliance"]
