"""Synthetic design domains (solid voxels, supports, loads, shell) of the benchmark
workloads.  Arrays are indexed [z, y, x]; vectors are SoA (3, nz+1, ny+1, nx+1).

cantilever(): box, all nodes at x = 0 fixed, unit total force in -z spread evenly over
the nodes of the bottom edge at x = max.  The reference applies the force field value
at every load node without normalisation (grid/Grid.cpp:2709-2716); here the total is
1, split evenly.

ellipsoids(): union of seeded random ellipsoids in a box, the connected part that
touches z = 0, solid nodes at z = 0 fixed, unit total load on the top-surface voxels
along the outward normal of the smoothed occupancy, 3-voxel shell.
"""
import numpy as np


def cantilever(nx, ny, nz):
    solid = np.ones((nz, ny, nx), np.uint8)
    support = np.zeros((nz + 1, ny + 1, nx + 1), np.uint8)
    support[:, :, 0] = 1
    f = np.zeros((3, nz + 1, ny + 1, nx + 1))
    nload = ny + 1
    f[2, 0, :, nx] = -1.0 / nload
    return solid, support, f.reshape(-1)


def embed_in_cube(solid, support=None, f=None, shell=None):
    """The reference's grid shape (padd2SatifyResolution, voxelizer/voxelize_main.cpp:
    312-340): the domain centred in the smallest power-of-two cube strictly larger than its
    longest side, at offset (side - n) // 2 per axis (24x12x12 -> 32^3 at (4, 10, 10)).
    Arrays are [z, y, x]; f is the flat (3, nz+1, ny+1, nx+1) load vector.  Returns
    (solid, support, f, shell, offset (x, y, z)) of the cube; absent inputs stay None."""
    nz, ny, nx = solid.shape
    side = 1
    while side <= max(nx, ny, nz):
        side *= 2
    ox, oy, oz = ((side - nx) // 2, (side - ny) // 2, (side - nz) // 2)
    S = np.zeros((side, side, side), solid.dtype)
    S[oz:oz + nz, oy:oy + ny, ox:ox + nx] = solid
    out = [S]
    if support is not None:
        U = np.zeros((side + 1,) * 3, support.dtype)
        U[oz:oz + nz + 1, oy:oy + ny + 1, ox:ox + nx + 1] = support
        out.append(U)
    else:
        out.append(None)
    if f is not None:
        F = np.zeros((3,) + (side + 1,) * 3)
        F[:, oz:oz + nz + 1, oy:oy + ny + 1, ox:ox + nx + 1] = np.asarray(f).reshape(3, nz + 1, ny + 1, nx + 1)
        out.append(F.reshape(-1))
    else:
        out.append(None)
    if shell is not None:
        H = np.zeros((side, side, side), shell.dtype)
        H[oz:oz + nz, oy:oy + ny, ox:ox + nx] = shell
        out.append(H)
    else:
        out.append(None)
    return (*out, (ox, oy, oz))


def face_load_region(nx, ny, nz, face="x+", inset=0):
    """Load region for the worst-case load search: the nodes of one box face ("x+",
    "y-", ...) at least `inset` nodes away from that face's border, with the face's
    outward unit normal each.  Returns (lexicographic node ids, (n, 3) normals)."""
    axis, sign = "xyz".index(face[0]), 1 if face[1] == "+" else -1
    dims = (nx + 1, ny + 1, nz + 1)
    grids = [np.arange(d) for d in dims]
    grids[axis] = np.array([dims[axis] - 1 if sign > 0 else 0])
    for a in range(3):
        if a != axis and inset:
            grids[a] = grids[a][inset:dims[a] - inset]
    z, y, x = np.meshgrid(grids[2], grids[1], grids[0], indexing="ij")
    ids = (x + dims[0] * (y + dims[1] * z)).ravel().astype(np.int64)
    nrm = np.zeros((ids.size, 3))
    nrm[:, axis] = sign
    return ids, nrm


def batched_patch_loads(nx, ny, nz, K=16, patch=8, seed=7):
    """K unit forces with directions uniform on the sphere, each spread over a seeded
    random patch x patch node patch of the x = max face."""
    rng = np.random.default_rng(seed)
    F = np.zeros((K, 3, nz + 1, ny + 1, nx + 1))
    dirs = []
    for k in range(K):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        oy = rng.integers(0, ny + 2 - patch)
        oz = rng.integers(0, nz + 2 - patch)
        for c in range(3):
            F[k, c, oz:oz + patch, oy:oy + patch, nx] = d[c] / (patch * patch)
        dirs.append(d)
    return F.reshape(K, -1), np.array(dirs)


def _ellipsoid_union(nx, ny, nz, cs, ax):
    solid = np.zeros((nz, ny, nx), bool)
    for (cx, cy, cz), (a, b, c) in zip(cs, ax):
        x0, x1 = max(int(cx - a) - 1, 0), min(int(cx + a) + 2, nx)
        y0, y1 = max(int(cy - b) - 1, 0), min(int(cy + b) + 2, ny)
        z0, z1 = max(int(cz - c) - 1, 0), min(int(cz + c) + 2, nz)
        if x0 >= x1 or y0 >= y1 or z0 >= z1:
            continue
        X2 = ((np.arange(x0, x1, dtype=np.float32) + 0.5 - cx) / a) ** 2
        Y2 = ((np.arange(y0, y1, dtype=np.float32) + 0.5 - cy) / b) ** 2
        Z2 = ((np.arange(z0, z1, dtype=np.float32) + 0.5 - cz) / c) ** 2
        xy = Y2[:, None] + X2[None, :]
        for k, z2 in enumerate(Z2):
            solid[z0 + k, y0:y1, x0:x1] |= xy <= (1.0 - z2)
    return solid


# The lattice ball {d : |d|^2 <= 9} is exactly the union of these cuboids (half-widths
# hx, hy, hz): every point of the ball lies in one, and each cuboid's far corner has
# |d|^2 = 9.
_BALL3 = ((0, 0, 3), (3, 0, 0), (0, 3, 0), (2, 1, 2), (1, 2, 2), (2, 2, 1))


def _within3_of_exterior(solid):
    """Solid voxels whose centre is within Euclidean distance 3 of a non-solid voxel
    centre inside the array (the same as distance_transform_edt(solid) <= 3)."""
    from scipy import ndimage
    ext = (~solid).astype(np.uint8)
    near = np.zeros(solid.shape, bool)
    for hx, hy, hz in _BALL3:
        d = ext
        for axis, hw in ((2, hx), (1, hy), (0, hz)):
            if hw:
                d = ndimage.maximum_filter1d(d, 2 * hw + 1, axis=axis, mode="constant", cval=0)
        near |= d.astype(bool)
    return solid & near


def ellipsoids(nx=511, ny=320, nz=320, n=6, seed=2024, target=0.35, shell_width=3,
               axes=(60, 180), max_tries=64):
    """Union of n axis-aligned ellipsoids (semi-axes ~U[axes], centres uniform in the
    central 60% of the box), the lowest one lowered to touch z = 0; keeps the part
    connected to z = 0 and the first draw whose fill is within 2% of `target`.
    Returns solid, support, loads, shell, meta."""
    from scipy import ndimage

    assert shell_width == 3, "the cuboid decomposition covers the radius-3 ball"
    best = None
    rng = np.random.default_rng(seed)
    for t in range(max_tries):
        cs = np.stack([rng.uniform(0.2 * nx, 0.8 * nx, n), rng.uniform(0.2 * ny, 0.8 * ny, n),
                       rng.uniform(0.2 * nz, 0.8 * nz, n)], 1)
        ax = rng.uniform(axes[0], axes[1], (n, 3))
        lo = np.argmin(cs[:, 2] - ax[:, 2])
        cs[lo, 2] = ax[lo, 2] * 0.8
        solid = _ellipsoid_union(nx, ny, nz, cs, ax)
        lab, nl = ndimage.label(solid)
        keep = np.zeros(nl + 1, bool)
        keep[np.unique(lab[0])] = True
        keep[0] = False
        conn = keep[lab]
        fill = conn.mean()
        if best is None or abs(fill - target) < abs(best[0] - target):
            best = (fill, conn, cs, ax, t)
        if abs(fill - target) <= 0.02:
            break
    fill, solid, cs, ax, tries = best
    solid_u8 = solid.astype(np.uint8)
    act = np.zeros((nz + 1, ny + 1, nx + 1), bool)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                act[dz:dz + nz, dy:dy + ny, dx:dx + nx] |= solid
    support = np.zeros_like(act)
    support[0] = act[0]
    # top-surface voxels: solid with an empty (or outside) +z neighbour
    top = solid.copy()
    top[:-1] &= ~solid[1:]
    occ = ndimage.uniform_filter(solid.astype(np.float32), size=3, mode="constant")
    tz, ty, tx = np.nonzero(top)
    f = np.zeros((3, nz + 1, ny + 1, nx + 1))
    if tz.size:
        # outward normal -grad(occ) by central differences at the top voxels
        def cd(axis, idx):
            hi = list(idx); lo_ = list(idx)
            hi[axis] = np.minimum(idx[axis] + 1, occ.shape[axis] - 1)
            lo_[axis] = np.maximum(idx[axis] - 1, 0)
            return (occ[tuple(hi)] - occ[tuple(lo_)]) * 0.5
        idx = (tz, ty, tx)
        v = -np.stack([cd(2, idx), cd(1, idx), cd(0, idx)], 0).astype(np.float64)
        ln = np.linalg.norm(v, axis=0)
        v = np.where(ln > 1e-12, v / np.maximum(ln, 1e-12), np.array([[0.0], [0.0], [1.0]]))
        wgt = 1.0 / (4 * tz.size)
        nn = (nz + 1) * (ny + 1) * (nx + 1)
        for (oy, ox) in ((0, 0), (0, 1), (1, 0), (1, 1)):
            flat = (tx + ox) + (nx + 1) * ((ty + oy) + (ny + 1) * (tz + 1))
            for c in range(3):
                f[c] += np.bincount(flat, weights=wgt * v[c], minlength=nn).reshape(f[c].shape)
    shell = _within3_of_exterior(solid).astype(np.uint8)
    meta = {"fill": float(fill), "tries": int(tries + 1), "centres": cs.tolist(),
            "semi_axes": ax.tolist(), "n_top": int(tz.size), "n_solid": int(solid.sum())}
    return solid_u8, support.astype(np.uint8), f.reshape(-1), shell, meta
