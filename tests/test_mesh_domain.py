"""Mesh-driven domains (topoptspline3d_amd.mesh_domain) on the reference's own bench input
bench/cube24x12x12 (cube24x12x12.obj + config11.json, copied as fixtures into
tests/golden/bench_cube24x12x12), and the pieces against independent constructions:
* the cube embedding of voxelizer/voxelize_main.cpp:312-340: the 24x12x12 box at
  resolution 24 sits in a 32^3 cube at voxel offset (4, 10, 10) with h = 1, and fills it;
* the voxeliser against the exact inside test of an octahedron |d|_1 < r whose edges
  project onto voxel-column centres (shared-edge ownership must count each column once);
* regions: the half-space side of config11's three points (x < 0.05), open / closed
  spheres, boxes, complements, cylinders, by hand;
* surface nodes of a 3^3 block: all but the 2^3 interior nodes;
* supports and loads of config11: the x = 0 face plus the two corner spheres; at
  resolution 48 the two load spheres hold one free node each, force (0, 0, -1);
* the worst-case loop runs on the config's domain over the oracle."""
import os

import numpy as np
import pytest

from oracle_api import ROOT, oracle_api
from topoptspline3d_amd import mesh_domain as M

CFG = os.path.join(ROOT, "tests", "golden", "bench_cube24x12x12", "config11.json")


def test_cube_embedding_and_voxels_of_the_bench_box():
    V, T = M.read_mesh(os.path.join(os.path.dirname(CFG), "cube24x12x12.obj"))
    R, corner, h = M.cube_embedding(V, 24)
    assert R == 32 and h == 1.0 and np.array_equal(corner, [-4, -10, -10])
    cube = M.voxelize(V, T, R, corner, h)
    ref = np.zeros((32, 32, 32), bool)
    ref[10:22, 10:22, 4:28] = True
    assert np.array_equal(cube, ref)
    R2, corner2, h2 = M.cube_embedding(V, 48)
    assert R2 == 64 and h2 == 0.5 and np.array_equal(corner2, [-4, -10, -10])


def octahedron(c, r):
    V = np.array([[r, 0, 0], [-r, 0, 0], [0, r, 0], [0, -r, 0], [0, 0, r], [0, 0, -r]], np.float64) + c
    T = []
    for sx, ix in ((1, 0), (-1, 1)):
        for sy, iy in ((1, 2), (-1, 3)):
            for sz, iz in ((1, 4), (-1, 5)):
                tri = [ix, iy, iz]
                if sx * sy * sz < 0:
                    tri = [ix, iz, iy]
                T.append(tri)
    return V.astype(np.float32), np.array(T)


@pytest.mark.parametrize("c", [(8.31, 8.13, 7.92), (8.5, 8.5, 8.75)])  # no centre on a face
def test_voxelizer_matches_inside_test(c):
    V, T = octahedron(np.array(c), 5.0)
    R, corner, h = 16, np.zeros(3, np.float32), 1.0
    cube = M.voxelize(V, T, R, corner, h)
    z, y, x = np.meshgrid(*(np.arange(16) + 0.5,) * 3, indexing="ij")
    ref = np.abs(x - c[0]) + np.abs(y - c[1]) + np.abs(z - c[2]) < 5.0
    assert np.array_equal(cube, ref), (cube ^ ref).sum()


def test_regions():
    X = np.array([[0.0, 0, 0], [0.05, 0, 0], [0.1, 0, 0], [1, 1, 1]])
    hs = M.Region({"type": "halfspace", "p1": [0.05, 0, 0], "p2": [0.05, 0, 1], "p3": [0.05, 1, 0]})
    assert hs.contains(X).tolist() == [True, False, False, False]
    hs_c = M.Region({"type": "halfspace", "p1": [0.05, 0, 0], "p2": [0.05, 0, 1], "p3": [0.05, 1, 0],
                     "closed": True})
    assert hs_c.contains(X).tolist() == [True, True, False, False]
    sp = M.Region({"type": "sphere", "center": [0, 0, 0], "radius": 0.1})
    assert sp.contains(X).tolist() == [True, True, False, False]
    assert M.Region({"type": "sphere", "center": [0, 0, 0], "radius": 0.1, "closed": True}).contains(X)[2]
    assert M.Region({"type": "sphere", "center": [0, 0, 0], "radius": 0.1, "complement": True}).contains(
        X).tolist() == [False, False, True, True]
    bx = M.Region({"type": "box", "p": [1.5, 1.5, 1.5], "q": [0.5, 0.5, 0.5]})
    assert bx.contains(X).tolist() == [False, False, False, True]
    cy = M.Region({"type": "cylinder", "bottomCenter": [1, 1, 0], "heightVector": [0, 0, 2], "radius": 0.5})
    assert cy.contains(np.array([[1, 1, 0], [1, 1, 2.1], [1.4, 1, 1], [1.5, 1, 1]])).tolist() == \
        [True, False, True, False]
    with pytest.raises(NotImplementedError):
        M.Region({"type": "torus"})


def test_surface_nodes_of_a_block():
    surf, act = M.surface_nodes(np.ones((3, 3, 3), np.uint8))
    assert act.all() and surf.sum() == 64 - 8 and not surf[1:3, 1:3, 1:3].any()
    s = np.zeros((3, 3, 3), np.uint8)
    s[1, 1, 1] = 1
    surf, act = M.surface_nodes(s)
    assert act.sum() == 8 and surf.sum() == 8


def test_config11_supports_and_loads():
    d = M.load_config(CFG, 24)
    assert d["solid"].shape == (12, 12, 24) and d["solid"].all()
    assert np.array_equal(d["cube"][1], [4, 10, 10])
    sup = d["support"]
    assert sup[:, :, 0].all() and sup.sum() == 13 * 13 + 2 and sup[0, 0, 24] and sup[0, 12, 24]
    assert d["load_nodes"].size == 0  # both load spheres only reach fixed x = 0 nodes at h = 1
    d = M.load_config(CFG, 48)
    assert d["h"] == 0.5 and sup.sum() == 171 and d["support"].sum() == 25 * 25 + 2
    nx1, ny1 = 49, 25
    pos = [(i % nx1 * 0.5, i // nx1 % ny1 * 0.5, i // (nx1 * ny1) * 0.5) for i in d["load_nodes"]]
    assert sorted(pos) == [(0.5, 0.0, 12.0), (0.5, 12.0, 12.0)]
    assert np.array_equal(d["load_force"], [[0, 0, -1], [0, 0, -1]])
    assert np.allclose(np.abs(d["normals"]), [[0, 2 ** -0.5, 2 ** -0.5]] * 2)
    f = d["f"].reshape(3, -1)
    assert f[2].sum() == -2 and np.count_nonzero(f) == 2


def test_shell_is_the_three_voxel_rind():
    """shell_width 3 on the bench box at resolution 48 (a 48x24x24 block): the elements
    within Euclidean distance 3 of the outside, i.e. all but the 42x18x18 core"""
    d = M.load_config(CFG, 48, shell_width=3)
    sh = d["shell"].astype(bool)
    core = np.zeros_like(sh)
    core[3:-3, 3:-3, 3:-3] = True
    assert np.array_equal(sh, ~core), (sh.sum(), (~core).sum())


def test_worst_case_loop_on_config11():
    from topoptspline3d_amd.drivers import run_worst_oc
    d = M.load_config(CFG, 48)
    rec = run_worst_oc(oracle_api(), 0, 0, 0, iters=2, domain=d)
    assert all(np.isfinite(rec["c"])) and all(c > 0 for c in rec["c"]) and max(rec["its"]) < 500
