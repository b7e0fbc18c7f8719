"""Pins the oracle's floating-body path (no supported node: the rigid motions are K's null
space) without the reference:
* forceProject (optimization/projection.cpp:289-350): on a floating body the projected load
  has zero resultant force and moment, equals the input minus its least-squares fit by
  the rigid motions of the load nodes (numpy lstsq), is idempotent, and with normals
  lies along them; with supports it is only restricted to the load nodes (and taken along
  the normals);
* the solve: with a balanced load the FCG converges, and after displacementProject
  (projection.cpp:374-395) the displacement equals pinv(K) f of the assembled stiffness
  (the minimum-norm solution, which the reference's BDCSVD coarsest solve targets);
* the worst-case search (modifiedPM with the floating projections): its fixed point is the
  top eigenpair of the balanced load-node block of pinv(K), free and along normals."""
import numpy as np
import pytest

from oracle_api import oracle_api
from test_oracle import E, NU, P, RMIN, global_K
from topoptspline3d_amd.context import Context, TosError
from topoptspline3d_amd.domains import cantilever, face_load_region


def floating(nx, ny, nz, seed=0, levels=2, api=None):
    solid = np.ones((nz, ny, nx), np.uint8)
    c = Context(api or oracle_api(), solid, None, n_levels=levels)
    c.set_material(E, NU, P, RMIN)
    rho = np.random.default_rng(seed).uniform(0.2, 1.0, solid.size).astype(np.float32)
    c.set_density(rho)
    c.set_loads(np.zeros(3 * c.nn))
    c.update_operator()
    return c, solid, c.density()


def coords(c, ids):
    nz, ny, nx = c.shape_n
    return np.stack([ids % nx, (ids // nx) % ny, ids // (nx * ny)], 1).astype(np.float64)


def rigid_matrix(p):
    """(3n, 6) rigid motions at points p (node-major rows): translations, rotations p^"""
    n = p.shape[0]
    R = np.zeros((3 * n, 6))
    for k in range(3):
        R[k::3, k] = 1.0
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    R[1::3, 3], R[2::3, 3] = z, -y
    R[0::3, 4], R[2::3, 4] = -z, x
    R[0::3, 5], R[1::3, 5] = y, -x
    return R


def region_force(c, ids):
    return c.get_loads(0).reshape(3, -1)[:, ids].T.ravel()


def set_region_force(c, ids, fs):
    f = np.zeros((3, c.nn))
    f[:, ids] = fs.reshape(-1, 3).T
    c.set_loads(f.ravel())


def resultant(p, fs):
    F = fs.reshape(-1, 3)
    return F.sum(0), np.cross(p - p.mean(0), F).sum(0)


def test_floating_flag():
    c, *_ = floating(4, 2, 2)
    assert c.info()["has_support"] is False
    solid, support, f = cantilever(4, 2, 2)
    s = Context(oracle_api(), solid, support)
    assert s.info()["has_support"] is True


@pytest.mark.parametrize("with_normals", [False, True])
def test_balance_loads_floating(with_normals):
    c, *_ = floating(8, 4, 4)
    ids, nrm = face_load_region(8, 4, 4, "x+")
    ids2, nrm2 = face_load_region(8, 4, 4, "z+", inset=1)
    ids, nrm = np.r_[ids, ids2], np.r_[nrm, nrm2]  # two faces: the moments are not degenerate
    rng = np.random.default_rng(1)
    fs = rng.normal(size=3 * ids.size)
    set_region_force(c, ids, fs)
    full_before = c.get_loads(0)
    c.balance_loads(ids, normals=nrm if with_normals else None)
    out = region_force(c, ids)
    full = c.get_loads(0).reshape(3, -1)
    mask = np.ones(c.nn, bool)
    mask[ids] = False
    assert np.all(full[:, mask] == 0)
    p = coords(c, ids)
    F, M = resultant(p, out)
    scale = np.abs(fs).max() * ids.size
    assert np.abs(F).max() <= 1e-12 * scale and np.abs(M).max() <= 1e-10 * scale
    R = rigid_matrix(p)
    if with_normals:
        N = np.zeros((3 * ids.size, ids.size))
        for i in range(ids.size):
            N[3 * i:3 * i + 3, i] = nrm[i]
        fn = N @ (N.T @ fs)
        B = N @ (N.T @ R)  # rigid motions along the normals
        ref = fn - B @ np.linalg.lstsq(B, fn, rcond=None)[0]
        side = np.cross(out.reshape(-1, 3), nrm)
        assert np.abs(side).max() <= 1e-12 * np.abs(fs).max()
    else:
        ref = fs - R @ np.linalg.lstsq(R, fs, rcond=None)[0]
    assert np.allclose(out, ref, rtol=0, atol=1e-11 * np.abs(fs).max())
    c.balance_loads(ids, normals=nrm if with_normals else None)  # idempotent
    assert np.allclose(region_force(c, ids), out, rtol=0, atol=1e-12 * np.abs(fs).max())
    assert not np.allclose(full_before, c.get_loads(0))


@pytest.mark.parametrize("with_normals", [False, True])
def test_balance_loads_with_supports_only_restricts(with_normals):
    solid, support, f = cantilever(8, 4, 4)
    c = Context(oracle_api(), solid, support)
    c.set_material(E, NU, P, RMIN)
    ids, nrm = face_load_region(8, 4, 4, "x+")
    rng = np.random.default_rng(2)
    full = rng.normal(size=(3, c.nn))
    full[:, (support.ravel() > 0)] = 0
    c.set_loads(full.ravel())
    c.balance_loads(ids, normals=nrm if with_normals else None)
    out = c.get_loads(0).reshape(3, -1)
    exp = np.zeros_like(full)
    exp[:, ids] = full[:, ids]
    if with_normals:
        exp[:, ids] = nrm.T * (nrm.T * full[:, ids]).sum(0)
    assert np.allclose(out, exp, rtol=0, atol=1e-15)


def test_remove_rigid_rejects_supported_context():
    solid, support, f = cantilever(4, 2, 2)
    c = Context(oracle_api(), solid, support)
    c.set_material(E, NU, P, RMIN)
    c.set_loads(f)
    with pytest.raises(TosError) as e:
        c.remove_rigid()
    assert e.value.code == -6


def test_floating_solve_is_the_minimum_norm_solution():
    nx, ny, nz = 8, 4, 4
    c, solid, rho = floating(nx, ny, nz, seed=3)
    assert c.info()["n_levels"] == 2 and c.info()["coarse_dofs"] > 0
    ids, _ = face_load_region(nx, ny, nz, "x+")
    ids0, _ = face_load_region(nx, ny, nz, "x-")
    f = np.zeros((3, c.nn))
    rng = np.random.default_rng(5)
    f[:, ids] = rng.normal(size=(3, ids.size))
    f[:, ids0] = rng.normal(size=(3, ids0.size))
    c.set_loads(f.ravel())
    both = np.r_[ids, ids0]
    c.balance_loads(both)
    fb = c.get_loads(0)
    st = c.solve(rtol=1e-11, maxit=400)
    assert st["rel_res"][0] <= 1e-11
    c.remove_rigid()
    u = c.displacement()
    K = global_K(c, solid, rho).toarray()
    u_ref = np.linalg.pinv(K, rcond=1e-10, hermitian=True) @ fb
    print(f"\n[floating] {st['iterations'][0]} FCG iterations; |u - pinv(K) f| / |u| = "
          f"{np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref):.2e}")
    assert np.linalg.norm(u - u_ref) <= 1e-8 * np.linalg.norm(u_ref)
    # and u carries no rigid motion over the active nodes
    allp = coords(c, np.arange(c.nn))
    ur = u.reshape(3, -1).T.ravel()
    assert np.abs(rigid_matrix(allp).T @ ur).max() <= 1e-10 * np.abs(ur).max() * c.nn


def test_unbalanced_load_does_not_converge():
    c, *_ = floating(8, 4, 4)
    ids, _ = face_load_region(8, 4, 4, "x+")
    f = np.zeros((3, c.nn))
    f[0, ids] = 1.0  # a net force: no equilibrium exists
    c.set_loads(f.ravel())
    st = c.solve(rtol=1e-8, maxit=60, check=False)
    assert st["rc"] == -3 and st["rel_res"][0] > 1e-3


@pytest.mark.parametrize("with_normals", [False, True])
def test_floating_worst_case_fixed_point_is_top_eigenpair(with_normals):
    nx, ny, nz = 8, 4, 4
    c, solid, rho = floating(nx, ny, nz, seed=6)
    ids, nrm = face_load_region(nx, ny, nz, "x+")
    ids2, nrm2 = face_load_region(nx, ny, nz, "x-")
    ids, nrm = np.r_[ids, ids2], np.r_[nrm, nrm2]
    K = global_K(c, solid, rho).toarray()
    Kp = np.linalg.pinv(K, rcond=1e-10, hermitian=True)
    cols = np.array([k * c.nn + n for n in ids for k in range(3)])
    S = Kp[np.ix_(cols, cols)]  # node-major load-node block
    R = rigid_matrix(coords(c, ids))
    if with_normals:
        N = np.zeros((3 * ids.size, ids.size))
        for i in range(ids.size):
            N[3 * i:3 * i + 3, i] = nrm[i]
        S, B = N.T @ S @ N, N.T @ R
    else:
        B = R
    U, sv, _ = np.linalg.svd(B, full_matrices=False)
    Q = U[:, sv > 1e-10 * sv.max()]  # range of the rigid motions (along normals: rank 3 here)
    Pb = np.eye(S.shape[0]) - Q @ Q.T  # balanced loads
    w, V = np.linalg.eigh(Pb @ (0.5 * (S + S.T)) @ Pb)
    lam, v = w[-1], V[:, -1]
    st = c.worst_compliance(ids, normals=nrm if with_normals else None, seed=3, max_iter=5000,
                            fch_tol=1e-9, res_tol=1e-12)
    print(f"\n[floating worst] normals={with_normals}: lambda_max {lam:.9e} (next {w[-2]:.6e}), "
          f"search {st['compliance']:.9e} in {st['iterations']} cycles")
    assert st["fch"] <= 1e-9 and st["rel_res"] <= 1e-12
    assert abs(st["compliance"] - lam) <= 1e-8 * lam
    fs = region_force(c, ids)
    fv = (fs.reshape(-1, 3) * nrm).sum(1) if with_normals else fs
    assert abs(abs(fv @ v) - 1.0) <= 1e-7
    F, M = resultant(coords(c, ids), fs)
    assert np.abs(F).max() <= 1e-12 and np.abs(M).max() <= 1e-10


def test_floating_solve_with_the_skip_layer():
    """a floating body on the skip-layer hierarchy: the coarsest factor's rigid motions sit
    on a level 8x coarser than the fine one, and the solve still returns pinv(K) f"""
    nx, ny, nz = 16, 8, 8
    solid = np.ones((nz, ny, nx), np.uint8)
    c = Context(oracle_api(), solid, None, n_levels=4, skip_layer=True)
    c.set_material(E, NU, P, RMIN)
    rho = np.random.default_rng(8).uniform(0.2, 1.0, solid.size).astype(np.float32)
    c.set_density(rho)
    assert c.info()["n_levels"] == 3 and c.info()["level_dims"][2] == (2, 1, 1)
    ids, _ = face_load_region(nx, ny, nz, "x+")
    ids0, _ = face_load_region(nx, ny, nz, "x-")
    both = np.r_[ids, ids0]
    f = np.zeros((3, c.nn))
    f[:, both] = np.random.default_rng(9).normal(size=(3, both.size))
    c.set_loads(f.ravel())
    c.balance_loads(both)
    c.update_operator()
    st = c.solve(rtol=1e-10, maxit=600)
    assert st["rel_res"][0] <= 1e-10
    c.remove_rigid()
    K = global_K(c, solid, c.density()).toarray()
    u_ref = np.linalg.pinv(K, rcond=1e-10, hermitian=True) @ c.get_loads(0)
    assert np.linalg.norm(c.displacement() - u_ref) <= 1e-8 * np.linalg.norm(u_ref)
