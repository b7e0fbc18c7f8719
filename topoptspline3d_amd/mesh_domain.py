"""Mesh-driven design domains: a triangle mesh and a JSON boundary-condition file (the
reference's bench/*/config*.json) -> voxel bits, supports, loads, load region, shell.

* Voxelisation (voxelizer/voxelize_main.cpp:292-406): the mesh's bounding box is padded to
  a cube of R = the smallest power of two above the requested resolution voxels, of edge
  h = (longest box side) / resolution, centred on the mesh (float32 arithmetic, as the
  reference's glm vectors); a voxel is solid when its centre is inside the mesh (parity of
  the mesh crossings of a +z ray below the centre, the solid voxelisation the reference
  calls).  The box is cropped to the solid voxels; node (i, j, k) of the crop sits at
  origin + (i, j, k) h with origin = cube corner + crop offset * h.
* Regions (config_parser.cpp:68-160, boundary/BoundaryCondition.h): "box" (corners p, q),
  "sphere" (center, radius), "halfspace" (p1, p2, p3: the side the normal
  (p2 - p1) x (p3 - p1) points to), "cylinder" (bottomCenter, heightVector, radius);
  "closed" adds the boundary, "complement" negates (not for cylinders).  A list of
  regions is their union.
* Force fields (config_parser.cpp:276-312): "const" = norm * direction inside its region;
  fields add up.  Other field types raise NotImplementedError.
* Boundary conditions (Grid::computeProjectionMatrix, grid/Grid.cpp:2660-2720): only
  surface nodes qualify (mark_surface_nodes_kernel, grid/Grid.cu:2781-2815: some axis
  along which no two solid elements straddle the node); a surface node in the fixed area
  is fixed; otherwise, in the load area, it is a load node with force = field(position)
  (no normalisation).  Load-node normals: the outward occupancy gradient over the node's
  eight elements (the reference takes them from the mesh with CGAL).
* Shell (shell_width > 0): elements within that Euclidean distance (in voxels) of the
  exterior have density 1 (the reference measures it to the mesh surface).
"""
import json
import os
import struct

import numpy as np


# ------------------------------------------------------------------------------------
def read_mesh(path):
    """(vertices (n, 3) float32, triangles (m, 3) int64) from an OBJ or STL file"""
    if path.lower().endswith(".stl"):
        return _read_stl(path)
    verts, tris = [], []
    with open(path, "r", errors="replace") as fh:
        for line in fh:
            tok = line.split()
            if not tok:
                continue
            if tok[0] == "v":
                verts.append([float(t) for t in tok[1:4]])
            elif tok[0] == "f":
                ids = [int(t.split("/")[0]) for t in tok[1:]]
                ids = [i - 1 if i > 0 else len(verts) + i for i in ids]
                for k in range(1, len(ids) - 1):  # fan triangulation of polygons
                    tris.append([ids[0], ids[k], ids[k + 1]])
    return np.array(verts, np.float32).reshape(-1, 3), np.array(tris, np.int64).reshape(-1, 3)


def _read_stl(path):
    raw = open(path, "rb").read()
    n = struct.unpack_from("<I", raw, 80)[0] if len(raw) >= 84 else -1
    if len(raw) == 84 + 50 * n:  # binary
        rec = np.frombuffer(raw, dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")]),
                            count=n, offset=84)
        v = rec["v"].reshape(-1, 3).astype(np.float32)
    else:
        v = np.array([[float(t) for t in l.split()[1:4]] for l in raw.decode(errors="replace").splitlines()
                      if l.strip().startswith("vertex")], np.float32)
    return v, np.arange(v.shape[0], dtype=np.int64).reshape(-1, 3)


def cube_embedding(verts, resolution):
    """(R, cube corner (3,) float32, h) of padd2SatifyResolution (voxelize_main.cpp:312-340)"""
    lo, hi = verts.min(0).astype(np.float32), verts.max(0).astype(np.float32)
    R = 1
    while R <= resolution:
        R <<= 1
    ln = hi - lo
    side = np.float32(ln.max()) * np.float32(R) / np.float32(resolution)
    corner = lo - (side - ln) / np.float32(2)
    return R, corner.astype(np.float32), float(side) / R


def voxelize(verts, tris, R, corner, h):
    """solid[z, y, x] (R^3, bool): voxel centres inside the mesh, by the parity of the
    crossings of a +z ray below each centre (edges shared by two triangles counted once:
    a column on an edge belongs to the triangle whose edge is a left or top edge)"""
    P = (verts.astype(np.float64) - corner.astype(np.float64)) / h  # voxel units
    flip = np.zeros((R, R, R), np.uint8)  # [z, y, x]: parity of the crossings starting there
    for t in tris:
        a, b, c = P[t[0]], P[t[1]], P[t[2]]
        area = (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])
        if area == 0:
            continue
        if area < 0:
            b, c = c, b
            area = -area
        x0, x1 = int(np.floor(min(a[0], b[0], c[0]) - 0.5)), int(np.ceil(max(a[0], b[0], c[0]) - 0.5))
        y0, y1 = int(np.floor(min(a[1], b[1], c[1]) - 0.5)), int(np.ceil(max(a[1], b[1], c[1]) - 0.5))
        x0, y0, x1, y1 = max(x0, 0), max(y0, 0), min(x1, R - 1), min(y1, R - 1)
        if x0 > x1 or y0 > y1:
            continue
        X, Y = np.meshgrid(np.arange(x0, x1 + 1) + 0.5, np.arange(y0, y1 + 1) + 0.5)
        inside = np.ones(X.shape, bool)
        w = []
        for p, q in ((b, c), (c, a), (a, b)):  # edge p -> q, weight of the opposite vertex
            e = (q[0] - p[0]) * (Y - p[1]) - (q[1] - p[1]) * (X - p[0])
            dy, dx = q[1] - p[1], q[0] - p[0]
            top_left = (dy < 0) or (dy == 0 and dx > 0)
            inside &= (e > 0) | ((e == 0) & top_left)
            w.append(e)
        if not inside.any():
            continue
        zhit = (w[0] * a[2] + w[1] * b[2] + w[2] * c[2]) / area
        iy, ix = np.nonzero(inside)
        zs = zhit[iy, ix]
        # crossings strictly below the centre z + 0.5: flip voxels z >= ceil(zhit - 0.5)
        z0 = np.clip(np.floor(zs - 0.5).astype(np.int64) + 1, 0, R)
        keep = z0 < R
        np.bitwise_xor.at(flip, (z0[keep], iy[keep] + y0, ix[keep] + x0), 1)
    return np.bitwise_xor.accumulate(flip, axis=0).astype(bool)


# ------------------------------------------------------------------------------------
class Region:
    def __init__(self, d):
        self.kind = d["type"]
        self.closed = bool(d.get("closed", False))
        self.complement = bool(d.get("complement", False)) and self.kind != "cylinder"
        if self.kind == "sphere":
            self.c, self.r = np.array(d["center"], float), float(d["radius"])
        elif self.kind == "box":
            p, q = np.array(d["p"], float), np.array(d["q"], float)
            self.lo, self.hi = np.minimum(p, q), np.maximum(p, q)
        elif self.kind == "halfspace":
            p1, p2, p3 = (np.array(d[k], float) for k in ("p1", "p2", "p3"))
            self.n, self.p = np.cross(p2 - p1, p3 - p1), p1
        elif self.kind == "cylinder":
            self.c, self.v, self.r = np.array(d["bottomCenter"], float), np.array(d["heightVector"], float), \
                float(d["radius"])
        else:
            raise NotImplementedError(f"region type {self.kind}")

    def contains(self, X):
        """X: (n, 3) positions -> (n,) bool"""
        eps = 1e-6  # boundary tolerance of "closed" regions
        if self.kind == "sphere":
            d2 = ((X - self.c) ** 2).sum(1)
            m = (d2 < self.r ** 2) | (self.closed & (np.abs(d2 - self.r ** 2) < eps))
        elif self.kind == "box":
            m = np.all((X > self.lo) & (X < self.hi), 1)
            if self.closed:
                m |= np.all((X >= self.lo - eps) & (X <= self.hi + eps), 1)
        elif self.kind == "halfspace":
            s = (X - self.p) @ self.n
            m = (s > 0) | (self.closed & (np.abs(s) <= eps))
        else:
            hgt = np.linalg.norm(self.v)
            z = self.v / hgt
            d = X - self.c
            proj = d @ z
            dev = np.linalg.norm(np.cross(d, z), axis=1)
            m = (proj <= hgt) & (proj >= 0) & (dev < self.r)
        return m ^ self.complement


def _union(regions, X):
    m = np.zeros(X.shape[0], bool)
    for r in regions:
        m |= r.contains(X)
    return m


def _field(fields, X):
    f = np.zeros((X.shape[0], 3))
    for region, vec in fields:
        f[region.contains(X)] += vec
    return f


def parse_config(path):
    cfg = json.load(open(path))
    fixed = [Region(d) for d in cfg.get("fixedarea", [])]
    load = [Region(d) for d in cfg.get("loadarea", [])]
    fields = []
    for d in cfg.get("force", []):
        if d.get("type") != "const":
            raise NotImplementedError(f"force field type {d.get('type')}")
        fields.append((Region(d["region"]), float(d["norm"]) * np.array(d["direction"], float)))
    return cfg, fixed, load, fields


# ------------------------------------------------------------------------------------
def surface_nodes(solid):
    """nodes[z, y, x] with some axis along which no two solid elements straddle the node"""
    s = np.pad(solid.astype(bool), 1)
    # e[dz][dy][dx]: element at node + (dx-1, dy-1, dz-1) solid
    nz, ny, nx = solid.shape
    e = [[[s[dz:dz + nz + 1, dy:dy + ny + 1, dx:dx + nx + 1] for dx in (0, 1)] for dy in (0, 1)] for dz in (0, 1)]
    ax = [np.zeros((nz + 1, ny + 1, nx + 1), bool) for _ in range(3)]
    for i in (0, 1):
        for j in (0, 1):
            ax[0] |= e[j][i][0] & e[j][i][1]
            ax[1] |= e[j][0][i] & e[j][1][i]
            ax[2] |= e[0][j][i] & e[1][j][i]
    active = np.zeros_like(ax[0])
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                active |= e[dz][dy][dx]
    return active & ~(ax[0] & ax[1] & ax[2]), active


def outward_normals(solid, nodes_zyx):
    """-(occupancy gradient) over the eight elements of each node, unit length (0 if flat)"""
    s = np.pad(solid.astype(np.float64), 1)
    z, y, x = nodes_zyx
    g = np.zeros((z.size, 3))
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                o = s[z + dz, y + dy, x + dx]
                g[:, 0] += o * (2 * dx - 1)
                g[:, 1] += o * (2 * dy - 1)
                g[:, 2] += o * (2 * dz - 1)
    n = np.linalg.norm(g, axis=1, keepdims=True)
    return np.where(n > 0, -g / np.where(n > 0, n, 1), 0.0)


def load_config(path, resolution, mesh_path=None, shell_width=0):
    """Design domain of a reference bench configuration at a grid resolution.  Returns a
    dict: solid[z, y, x] (uint8), support (node mask), f (SoA 3 x nodes), shell (or None),
    load_nodes (lexicographic ids), normals ((n, 3)), origin, h, cube (R, offset)."""
    cfg, fixed, load, fields = parse_config(path)
    mesh_path = mesh_path or os.path.join(os.path.dirname(path), cfg["model"])
    V, T = read_mesh(mesh_path)
    R, corner, h = cube_embedding(V, resolution)
    cube = voxelize(V, T, R, corner, h)
    zz, yy, xx = np.nonzero(cube)
    if zz.size == 0:
        raise ValueError("the mesh encloses no voxel centre at this resolution")
    o = np.array([xx.min(), yy.min(), zz.min()])
    solid = cube[zz.min():zz.max() + 1, yy.min():yy.max() + 1, xx.min():xx.max() + 1].astype(np.uint8)
    origin = corner.astype(np.float64) + o * h
    nz, ny, nx = solid.shape
    surf, active = surface_nodes(solid)
    Z, Y, X = np.nonzero(surf)
    pos = origin + np.stack([X, Y, Z], 1) * h
    is_fix = _union(fixed, pos)
    is_load = ~is_fix & _union(load, pos)
    support = np.zeros((nz + 1, ny + 1, nx + 1), np.uint8)
    support[Z[is_fix], Y[is_fix], X[is_fix]] = 1
    lz, ly, lx = Z[is_load], Y[is_load], X[is_load]
    ids = (lx + (nx + 1) * (ly + (ny + 1) * lz)).astype(np.int64)
    fl = _field(fields, pos[is_load])
    f = np.zeros((3, (nx + 1) * (ny + 1) * (nz + 1)))
    f[:, ids] = fl.T
    shell = None
    if shell_width:
        from .domains import _within3_of_exterior
        if shell_width != 3:
            raise NotImplementedError("shell widths other than 3 voxels")
        # one void voxel around the tight crop, so the exterior next to every face counts
        shell = _within3_of_exterior(np.pad(solid > 0, 1))[1:-1, 1:-1, 1:-1].astype(np.uint8)
    return dict(solid=solid, support=support, f=f.ravel(), shell=shell, load_nodes=ids,
                normals=outward_normals(solid, (lz, ly, lx)), load_force=fl, origin=origin, h=h,
                cube=(R, o), config=cfg)
