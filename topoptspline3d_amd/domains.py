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


def ellipsoids(nx=511, ny=320, nz=320, n=6, seed=2024, target=0.35, shell_width=3,
               axes=(60, 180), max_tries=64):
    """Union of n axis-aligned ellipsoids (semi-axes ~U[axes], centres uniform in the
    central 60% of the box) whose fill is closest to `target`; keeps the component
    connected to z = 0.  Returns solid, support, loads, shell, meta."""
    from scipy import ndimage

    best = None
    rng = np.random.default_rng(seed)
    z = np.arange(nz, dtype=np.float32)[:, None, None] + 0.5
    y = np.arange(ny, dtype=np.float32)[None, :, None] + 0.5
    x = np.arange(nx, dtype=np.float32)[None, None, :] + 0.5
    for t in range(max_tries):
        cs = np.stack([rng.uniform(0.2 * nx, 0.8 * nx, n), rng.uniform(0.2 * ny, 0.8 * ny, n),
                       rng.uniform(0.2 * nz, 0.8 * nz, n)], 1)
        ax = rng.uniform(axes[0], axes[1], (n, 3))
        # the lowest ellipsoid is dropped onto z = 0 so that the solid touches the support
        lo = np.argmin(cs[:, 2] - ax[:, 2])
        cs[lo, 2] = ax[lo, 2] * 0.8
        solid = np.zeros((nz, ny, nx), bool)
        for (cx, cy, cz), (a, b, c) in zip(cs, ax):
            solid |= ((x - cx) / a) ** 2 + ((y - cy) / b) ** 2 + ((z - cz) / c) ** 2 <= 1.0
        lab, nl = ndimage.label(solid)
        keep = np.unique(lab[0][lab[0] > 0])
        conn = np.isin(lab, keep)
        fill = conn.mean()
        if best is None or abs(fill - target) < abs(best[0] - target):
            best = (fill, conn, cs, ax, t)
        if abs(fill - target) <= 0.02:
            break
    fill, solid, cs, ax, tries = best
    solid_u8 = solid.astype(np.uint8)
    # nodes of active elements
    act = np.zeros((nz + 1, ny + 1, nx + 1), bool)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                act[dz:dz + nz, dy:dy + ny, dx:dx + nx] |= solid
    support = np.zeros_like(act)
    support[0] = act[0]
    # top-surface voxels: solid with empty (or outside) +z neighbour
    top = solid.copy()
    top[:-1] &= ~solid[1:]
    occ = ndimage.uniform_filter(solid.astype(np.float32), size=3, mode="constant")
    gz, gy, gx = np.gradient(occ)
    nrm = -np.stack([gx, gy, gz], 0)
    f = np.zeros((3, nz + 1, ny + 1, nx + 1))
    tz, ty, tx = np.nonzero(top)
    if tz.size:
        v = nrm[:, tz, ty, tx]
        ln = np.linalg.norm(v, axis=0)
        v = np.where(ln > 1e-12, v / np.maximum(ln, 1e-12), np.array([[0], [0], [1.0]]))
        wgt = 1.0 / (4 * tz.size)
        for (oy, ox) in ((0, 0), (0, 1), (1, 0), (1, 1)):
            for c in range(3):
                np.add.at(f[c], (tz + 1, ty + oy, tx + ox), wgt * v[c])
    # shell: elements within shell_width of the exterior
    dist = ndimage.distance_transform_edt(solid)
    shell = (solid & (dist <= shell_width)).astype(np.uint8)
    meta = {"fill": float(fill), "tries": int(tries + 1), "centres": cs.tolist(),
            "semi_axes": ax.tolist(), "n_top": int(tz.size), "n_solid": int(solid.sum())}
    return solid_u8, support.astype(np.uint8), f.reshape(-1), shell, meta
