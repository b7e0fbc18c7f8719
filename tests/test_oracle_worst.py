"""Pins the oracle's worst-case load search (modified power method, optimization/
optimization.cpp:213-351) without the reference:
* the start force: splitmix64's published outputs for seed 0 (Steele, Lea, Flood 2014),
  mapped to [-1, 1) and scaled to unit norm;
* the fixed point: with tight stop tolerances the search ends at the largest eigenvalue
  of the load-node block of K^-1 (free direction) or of N^T K^-1 N (force along the node
  normals N), computed here with scipy from the assembled stiffness matrix, and its force
  is that eigenvector;
* the reference's own stop rule (force change 1e-4, cycle residual 1e-2) ends within
  1e-4 of that eigenvalue on these cases;
* error paths: repeated, supported or out-of-range load nodes, zero normals."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle_api import oracle_api
from test_oracle import E, NU, P, RMIN, free_mask, global_K
from topoptspline3d_amd.context import Context, TosError
from topoptspline3d_amd.domains import cantilever, face_load_region

SPLITMIX0 = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def make(nx, ny, nz, seed=0, levels=2):
    solid, support, f = cantilever(nx, ny, nz)
    c = Context(oracle_api(), solid, support, n_levels=levels)
    c.set_material(E, NU, P, RMIN)
    c.set_loads(f)
    rho = np.random.default_rng(seed).uniform(0.2, 1.0, solid.size).astype(np.float32)
    c.set_density(rho)
    c.update_operator()
    return c, solid, c.density()


def load_block_inverse(c, solid, rho, nodes):
    """the load-node block of K^-1 (3n x 3n, rows/cols node-major: 3i + component)"""
    K = global_K(c, solid, rho)
    idx = np.nonzero(np.tile(free_mask(c), 3))[0]
    Kf = K[idx][:, idx].tocsc()
    pos = {d: j for j, d in enumerate(idx)}
    cols = [pos[k * c.nn + n] for n in nodes for k in range(3)]
    rhs = np.zeros((idx.size, len(cols)))
    rhs[cols, np.arange(len(cols))] = 1.0
    X = spla.splu(Kf).solve(rhs)
    return X[cols, :]


def force_on(c, nodes):
    f = c.get_loads(0).reshape(3, -1)
    return f[:, nodes].T.ravel()  # node-major


def test_start_force_is_splitmix64():
    c, *_ = make(6, 3, 4)
    ids, _ = face_load_region(6, 3, 4, "x+")
    st = c.worst_compliance(ids[:1], seed=0, max_iter=0)
    assert st["iterations"] == 0
    g = np.array([(z >> 11) * 2.0 ** -52 - 1.0 for z in SPLITMIX0])
    f = force_on(c, ids[:1])
    assert np.allclose(f, g / np.linalg.norm(g), rtol=0, atol=1e-15)
    full = c.get_loads(0)
    assert np.count_nonzero(full) == 3 and np.allclose(np.linalg.norm(full), 1.0, rtol=1e-15)


@pytest.mark.parametrize("with_normals", [False, True])
def test_fixed_point_is_top_eigenpair(with_normals):
    nx, ny, nz = 12, 4, 6
    c, solid, rho = make(nx, ny, nz, seed=4)
    ids, nrm = face_load_region(nx, ny, nz, "x+")
    S = load_block_inverse(c, solid, rho, ids)
    if with_normals:
        N = np.zeros((3 * ids.size, ids.size))
        for i in range(ids.size):
            N[3 * i:3 * i + 3, i] = nrm[i]
        S = N.T @ S @ N
    w, V = np.linalg.eigh(0.5 * (S + S.T))
    lam, v = w[-1], V[:, -1]
    print(f"\n[worst oracle] normals={with_normals} lambda_max {lam:.6e} (next {w[-2]:.6e})")
    st = c.worst_compliance(ids, normals=nrm if with_normals else None, seed=3, max_iter=5000,
                            fch_tol=1e-9, res_tol=1e-12)
    print(f"[worst oracle] iterations {st['iterations']} compliance {st['compliance']:.9e}")
    assert st["fch"] <= 1e-9 and st["rel_res"] <= 1e-12
    assert abs(st["compliance"] - lam) <= 1e-9 * lam
    f = force_on(c, ids)
    fv = (f.reshape(-1, 3) * nrm).sum(1) if with_normals else f  # coordinates of S's basis
    assert abs(abs(fv @ v) - 1.0) <= 1e-8
    assert np.allclose(c.compliance()[0], st["compliance"], rtol=1e-15)
    # the reference's stop rule lands close to the same eigenvalue
    st2 = c.worst_compliance(ids, normals=nrm if with_normals else None, seed=3)
    print(f"[worst oracle] default stop: iterations {st2['iterations']} compliance {st2['compliance']:.9e} "
          f"fch {st2['fch']:.2e} res {st2['rel_res']:.2e}")
    assert st2["fch"] <= 1e-4 and st2["rel_res"] <= 1e-2
    assert abs(st2["compliance"] - lam) <= 1e-4 * lam


def test_bad_load_regions_rejected():
    c, *_ = make(4, 3, 3)
    ids, nrm = face_load_region(4, 3, 3, "x+")
    for nodes, normals in ((np.r_[ids, ids[:1]], None),        # repeated
                           (np.array([0]), None),             # supported (x = 0)
                           (np.array([c.nn]), None),          # out of range
                           (ids, np.zeros((ids.size, 3)))):   # zero normal
        with pytest.raises(TosError) as e:
            c.worst_compliance(nodes, normals=normals)
        assert e.value.code == -1
