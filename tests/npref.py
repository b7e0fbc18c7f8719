"""Independent numpy/scipy constructions used to pin the oracle on tiny grids.

Nothing here reuses the oracle's code: the element matrix is integrated with a 3x3x3
Gauss rule from shape functions written out separately, the operator is an assembled
scipy sparse matrix, the prolongation an explicit sparse trilinear interpolation, and
the V-cycle runs on those matrices.
"""
import itertools

import numpy as np
import scipy.sparse as sp


def ke_gauss3(E=1e6, nu=0.4, h=1.0):
    """Element stiffness of the 8-node trilinear hex by 3x3x3 Gauss-Legendre quadrature
    (exact for this polynomial degree, so it must equal the 2x2x2 result)."""
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[np.arange(3), np.arange(3)] += 2 * mu
    D[3:, 3:] = np.eye(3) * mu
    pts, wts = np.polynomial.legendre.leggauss(3)
    pts = (pts + 1) / 2 * h
    wts = wts / 2 * h
    corners = [(i & 1, (i >> 1) & 1, i >> 2) for i in range(8)]
    K = np.zeros((24, 24))
    for (x, wx), (y, wy), (z, wz) in itertools.product(zip(pts, wts), repeat=3):
        grads = []
        for (a, b, c) in corners:
            fx = x / h if a else 1 - x / h
            fy = y / h if b else 1 - y / h
            fz = z / h if c else 1 - z / h
            dfx = (1 if a else -1) / h
            dfy = (1 if b else -1) / h
            dfz = (1 if c else -1) / h
            grads.append((dfx * fy * fz, fx * dfy * fz, fx * fy * dfz))
        B = np.zeros((6, 24))
        for n, (gx, gy, gz) in enumerate(grads):
            B[0, 3 * n] = gx
            B[1, 3 * n + 1] = gy
            B[2, 3 * n + 2] = gz
            B[3, 3 * n + 1] = gz; B[3, 3 * n + 2] = gy
            B[4, 3 * n + 0] = gz; B[4, 3 * n + 2] = gx
            B[5, 3 * n + 0] = gy; B[5, 3 * n + 1] = gx
        K += B.T @ D @ B * wx * wy * wz
    return K


def rigid_modes_element(h=1.0):
    R = np.zeros((24, 6))
    for i in range(8):
        p = np.array([i & 1, (i >> 1) & 1, i >> 2], float) * h
        R[3 * i:3 * i + 3, :3] = np.eye(3)
        R[3 * i:3 * i + 3, 3:] = np.array([[0, p[2], -p[1]], [-p[2], 0, p[0]], [p[1], -p[0], 0]])
    return R


def assemble(Ke, w, nx, ny, nz):
    """Global stiffness over all (dense) nodes, DOF order = component-major SoA
    (c*nn + node), element weights w[e] (0 = void)."""
    nn = (nx + 1) * (ny + 1) * (nz + 1)
    rows, cols, vals = [], [], []
    for ez in range(nz):
        for ey in range(ny):
            for ex in range(nx):
                e = ex + nx * (ey + ny * ez)
                if w[e] == 0:
                    continue
                nodes = [(ex + (i & 1)) + (nx + 1) * ((ey + ((i >> 1) & 1)) + (ny + 1) * (ez + (i >> 2)))
                         for i in range(8)]
                dofs = np.array([c * nn + n for n in nodes for c in range(3)])
                rows.append(np.repeat(dofs, 24))
                cols.append(np.tile(dofs, 24))
                vals.append((w[e] * Ke).ravel())
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(3 * nn, 3 * nn))


def prolongation(fdims, cdims):
    """Trilinear interpolation from coarse nodes (node dims cdims) to fine nodes (fdims),
    scalar, node ids x fastest."""
    fx, fy, fz = fdims
    cx, cy, cz = cdims
    rows, cols, vals = [], [], []
    for z in range(fz):
        for y in range(fy):
            for x in range(fx):
                a = x + fx * (y + fy * z)
                opts = []
                for v in (x, y, z):
                    opts.append([(v // 2, 1.0)] if v % 2 == 0 else [(v // 2, 0.5), (v // 2 + 1, 0.5)])
                for (X, wx), (Y, wy), (Z, wz) in itertools.product(*opts):
                    rows.append(a); cols.append(X + cx * (Y + cy * Z)); vals.append(wx * wy * wz)
    return sp.csr_matrix((vals, (rows, cols)), shape=(fx * fy * fz, cx * cy * cz))


def vec3(P):
    return sp.block_diag([P, P, P]).tocsr()


def colour_gs(A, u, f, free, ndims, backward):
    """8-colour nodal block GS with component-wise relaxation, on an assembled matrix."""
    nxn, nyn, nzn = ndims
    nn = nxn * nyn * nzn
    A = A.tocsr()
    order = range(7, -1, -1) if backward else range(8)
    for col in order:
        px, py, pz = col & 1, (col >> 1) & 1, col >> 2
        nodes = [x + nxn * (y + nyn * z)
                 for z in range(pz, nzn, 2) for y in range(py, nyn, 2) for x in range(px, nxn, 2)]
        unew = u.copy()
        for i in nodes:
            if not free[i]:
                continue
            comps = (2, 1, 0) if backward else (0, 1, 2)
            v = u[[i, nn + i, 2 * nn + i]].copy()
            for c in comps:
                row = A.getrow(c * nn + i)
                s = row.dot(u)[0] - sum(A[c * nn + i, q * nn + i] * u[q * nn + i] for q in range(3))
                others = sum(A[c * nn + i, q * nn + i] * v[q] for q in range(3) if q != c)
                v[c] = (f[c * nn + i] - s - others) / A[c * nn + i, c * nn + i]
            unew[[i, nn + i, 2 * nn + i]] = v
        u = unew
    return u
