"""Pins the oracle's skip-layer hierarchy (set_skip_layer, optimization/optimization.cpp:20)
without the reference:
* the hierarchy: dyadic level 1 dropped, the first coarse level ceil(n/4) elements per
  axis, every node of its active elements an unknown;
* its Galerkin operator equals P4^T K~ P4 built here with scipy: P4 the factor-4 trilinear
  interpolation (weights prod (4 - |d|)/4), K~ the fine stiffness with every supported
  node's rows and columns dropped and 1e6 per incident active element on its diagonal
  (restrict_stencil_nondyadic_OTFA_WS_kernel grid/Grid.cu:1112-1225);
* the V-cycle equals a scipy V-cycle over those operators with factor-4 restriction and
  prolongation (grid/Grid.cu:2509-2570, 2635-2671) between levels 0 and 1 and dyadic ones
  below, and is symmetric;
* solves with and without the skip reach the same displacement (the hierarchy only
  changes the preconditioner)."""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp

import npref
from oracle_api import oracle_api
from test_oracle import E, NU, P, RMIN, free_mask, global_K
from topoptspline3d_amd.context import Context
from topoptspline3d_amd.domains import cantilever


def make(nx, ny, nz, levels, skip=True, seed=3, void_frac=0.2, api=None):
    rng = np.random.default_rng(seed)
    solid, support, f = cantilever(nx, ny, nz)
    if void_frac:
        solid = (rng.random(solid.shape) > void_frac).astype(np.uint8)
        solid[:, :, 0] = 1
        solid[0, :, -1] = 1
    c = Context(api or oracle_api(), solid, support, n_levels=levels, skip_layer=skip)
    c.set_material(E, NU, P, RMIN)
    c.set_loads(f)
    c.set_density(rng.uniform(0.05, 1.0, solid.size).astype(np.float32))
    c.update_operator()
    return c, solid, support, f, c.density()


def p4(fdims, cdims):
    """factor-4 trilinear interpolation (scalar) from coarse nodes to fine nodes"""
    fx, fy, fz = fdims
    cx, cy, cz = cdims
    rows, cols, vals = [], [], []
    for z in range(fz):
        for y in range(fy):
            for x in range(fx):
                a = x + fx * (y + fy * z)
                opts = []
                for v in (x, y, z):
                    t = v % 4
                    opts.append([(v // 4, 1.0)] if t == 0 else [(v // 4, (4 - t) / 4), (v // 4 + 1, t / 4)])
                for (X, wx), (Y, wy), (Z, wz) in itertools.product(*opts):
                    rows.append(a); cols.append(X + cx * (Y + cy * Z)); vals.append(wx * wy * wz)
    return sp.csr_matrix((vals, (rows, cols)), shape=(fx * fy * fz, cx * cy * cz))


def penalised_K(solid, support, rho):
    """fine stiffness over all nodes with the skip-layer Dirichlet penalty"""
    nz, ny, nx = solid.shape
    Ke = npref.ke_gauss3(E, NU, 1.0)
    nn = (nx + 1) * (ny + 1) * (nz + 1)
    sup = support.ravel() > 0
    rows, cols, vals = [], [], []
    for ez, ey, ex in itertools.product(range(nz), range(ny), range(nx)):
        e = ex + nx * (ey + ny * ez)
        if not solid.ravel()[e]:
            continue
        nodes = [(ex + (i & 1)) + (nx + 1) * ((ey + ((i >> 1) & 1)) + (ny + 1) * (ez + (i >> 2))) for i in range(8)]
        M = float(rho[e]) ** P * Ke
        fixed = np.repeat([sup[n] for n in nodes], 3)
        M[fixed, :] = 0.0
        M[:, fixed] = 0.0
        M[np.nonzero(fixed)[0], np.nonzero(fixed)[0]] = 1e6
        dofs = np.array([c * nn + n for n in nodes for c in range(3)])
        rows.append(np.repeat(dofs, 24)); cols.append(np.tile(dofs, 24)); vals.append(M.ravel())
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(3 * nn, 3 * nn))


def stencil_matrix(st, dims):
    nxn, nyn, nzn = dims
    nn = nxn * nyn * nzn
    A = np.zeros((3 * nn, 3 * nn))
    for z, y, x in itertools.product(range(nzn), range(nyn), range(nxn)):
        i = x + nxn * (y + nyn * z)
        for k in range(27):
            jx, jy, jz = x + k % 3 - 1, y + (k // 3) % 3 - 1, z + k // 9 - 1
            if not (0 <= jx < nxn and 0 <= jy < nyn and 0 <= jz < nzn):
                continue
            j = jx + nxn * (jy + nyn * jz)
            for r in range(3):
                for s in range(3):
                    A[r * nn + i, s * nn + j] = st[k, r * 3 + s, i]
    return A


def test_skip_hierarchy():
    c, *_ = make(16, 8, 8, levels=4, void_frac=0.0)
    d = make(16, 8, 8, levels=4, skip=False, void_frac=0.0)[0]
    ic, idd = c.info(), d.info()
    assert idd["n_levels"] == 4 and ic["n_levels"] == 3
    assert ic["level_dims"] == [(16, 8, 8), (4, 2, 2), (2, 1, 1)]
    assert ic["n_free_nodes"][1] == 5 * 3 * 3  # every node of the active coarse elements
    # too shallow: the skip is ignored
    s2 = make(8, 4, 4, levels=2, void_frac=0.0)[0]
    assert s2.info()["level_dims"] == [(8, 4, 4), (4, 2, 2)]


def test_skip_galerkin_is_the_penalised_triple_product():
    c, solid, support, f, rho = make(12, 8, 8, levels=3)
    info = c.info()
    dims = [tuple(d + 1 for d in ld) for ld in info["level_dims"]]
    assert info["level_dims"][1] == (3, 2, 2)
    Kt = penalised_K(solid, support, rho)
    Pm = npref.vec3(p4(dims[0], dims[1]))
    ref = (Pm.T @ Kt @ Pm).toarray()
    fr1 = np.tile((c.node_flags(1) & 4) > 0, 3)
    ref[~fr1, :] = 0
    ref[:, ~fr1] = 0
    got = stencil_matrix(c.stencil(1), dims[1])
    scale = np.abs(ref).max()
    print(f"\n[skip galerkin] max |diff| / max = {np.abs(got - ref).max() / scale:.2e}")
    assert np.abs(got - ref).max() <= 1e-12 * scale


def test_skip_vcycle_matches_sparse_reference_and_is_symmetric():
    c, solid, support, f, rho = make(16, 8, 8, levels=4, seed=5, void_frac=0.15)
    info = c.info()
    assert info["n_levels"] == 3
    dims = [tuple(d + 1 for d in ld) for ld in info["level_dims"]]
    frees = [(c.node_flags(l) & 4) > 0 for l in range(3)]
    K = global_K(c, solid, rho)
    D0 = sp.diags(np.tile(frees[0], 3).astype(float))
    Ps = [npref.vec3(p4(dims[0], dims[1])).tocsr(),
          sp.diags(np.tile(frees[1], 3).astype(float)) @ npref.vec3(npref.prolongation(dims[1], dims[2]))]
    D1 = sp.diags(np.tile(frees[1], 3).astype(float))
    As = [(D0 @ K @ D0).tocsr(), (D1 @ (Ps[0].T @ penalised_K(solid, support, rho) @ Ps[0]) @ D1).tocsr()]
    As.append((Ps[1].T @ As[1] @ Ps[1]).tocsr())

    def vc(l, fvec):
        fr3 = np.tile(frees[l], 3)
        if l == 2:
            idx = np.nonzero(fr3)[0]
            u = np.zeros_like(fvec)
            u[idx] = np.linalg.solve(As[l][idx][:, idx].toarray(), fvec[idx])
            return u
        u = npref.colour_gs(As[l], np.zeros_like(fvec), fvec, frees[l], dims[l], backward=False)
        r = np.where(fr3, fvec - As[l] @ u, 0)
        fc = np.where(np.tile(frees[l + 1], 3), Ps[l].T @ r, 0)
        u = u + np.where(fr3, Ps[l] @ vc(l + 1, fc), 0)
        return npref.colour_gs(As[l], u, fvec, frees[l], dims[l], backward=True)

    rng = np.random.default_rng(9)
    fr3 = np.tile(frees[0], 3)
    r1 = np.where(fr3, rng.normal(size=3 * c.nn), 0)
    r2 = np.where(fr3, rng.normal(size=3 * c.nn), 0)
    z1 = c.vcycle(r1)
    ref = vc(0, r1)
    print(f"\n[skip vcycle] max |diff| / max = {np.abs(z1 - ref).max() / np.abs(ref).max():.2e}")
    assert np.allclose(z1, ref, rtol=0, atol=1e-10 * np.abs(ref).max())
    z2 = c.vcycle(r2)
    assert abs(z1 @ r2 - z2 @ r1) <= 1e-10 * abs(z1 @ r2)
    assert z1 @ r1 > 0


def test_skip_solve_matches_dyadic():
    a = make(32, 16, 16, levels=0, seed=7)[0]
    b = make(32, 16, 16, levels=0, seed=7, skip=False)[0]
    assert a.info()["n_levels"] == b.info()["n_levels"] - 1
    sa = a.solve(rtol=1e-10, maxit=400)
    sb = b.solve(rtol=1e-10, maxit=400)
    print(f"\n[skip solve] FCG: skip {sa['iterations'][0]} iterations, dyadic {sb['iterations'][0]}")
    assert sa["rel_res"][0] <= 1e-10 and sb["rel_res"][0] <= 1e-10
    ua, ub = a.displacement(), b.displacement()
    assert np.linalg.norm(ua - ub) <= 1e-8 * np.linalg.norm(ub)


def test_skip_vcycle_iteration_meets_the_reference_stop_rule():
    """the reference's solver: plain V(1,1) cycles until the relative residual is below
    1e-3, at most 50 (utils/test_utils.cpp:659-663), on a uniform density"""
    a = make(32, 16, 16, levels=0, seed=7, void_frac=0.0)[0]
    b = make(32, 16, 16, levels=0, seed=7, void_frac=0.0, skip=False)[0]
    for c in (a, b):
        c.set_density(np.full(32 * 16 * 16, 0.5, np.float32))
        c.update_operator()
    sa = a.solve(rtol=1e-3, maxit=50, use_pcg=False)
    sb = b.solve(rtol=1e-10, maxit=200)
    print(f"\n[skip V-cycles] {sa['iterations'][0]} cycles to {sa['rel_res'][0]:.2e}")
    assert sa["rel_res"][0] <= 1e-3 and sa["iterations"][0] <= 50
    ua, ub = a.displacement(), b.displacement()
    assert np.linalg.norm(ua - ub) <= 2e-2 * np.linalg.norm(ub)


def test_reference_shaped_c0():
    """24x12x12 cantilever in the reference's 32^3 cube (offset (4, 10, 10)) with the
    skip layer: the first coarse level is 6x4x4 active elements with 175 nodes, as the
    reference's factor-4 layer 2 straddles its cells (voxelize_main.cpp:331-335,
    Grid.cu:4453-4473); an ordinary SIMP step on it agrees with the cropped dyadic box."""
    from topoptspline3d_amd.domains import embed_in_cube
    solid, support, f = cantilever(24, 12, 12)
    S, U, F, _, off = embed_in_cube(solid, support, f)
    assert S.shape == (32, 32, 32) and off == (4, 10, 10)
    c = Context(oracle_api(), S, U, skip_layer=True)
    c.set_material(E, NU, P, RMIN)
    c.set_loads(F)
    c.set_density(np.where(S.ravel() > 0, 0.4, 0).astype(np.float32))
    c.update_operator()
    info = c.info()
    assert info["n_levels"] == 2 and info["n_active_elements"][1] == 96 and info["n_free_nodes"][1] == 175
    st = c.solve(rtol=1e-10, maxit=300)
    d = Context(oracle_api(), solid, support)
    d.set_material(E, NU, P, RMIN)
    d.set_loads(f)
    d.set_density(np.full(solid.size, 0.4, np.float32))
    d.update_operator()
    d.solve(rtol=1e-10, maxit=300)
    assert abs(c.compliance()[0] - d.compliance()[0]) <= 1e-9 * d.compliance()[0]
    ox, oy, oz = off
    u = c.displacement().reshape(3, 33, 33, 33)[:, oz:oz + 13, oy:oy + 13, ox:ox + 25]
    assert np.allclose(u.ravel(), d.displacement(), rtol=0, atol=1e-8 * np.abs(d.displacement()).max())
