"""Pins of the oracle's self-supporting and drip constraints on spline surface points
(deal_surface_points, optimization/optimization.cpp:1103-1128; grid/Grid.cu:5529-7156):
a hand-computed plane, and finite differences of every constraint mode's coefficient
gradient with the surface points held fixed (as the reference's gradients are)."""
import ctypes as C

import numpy as np
import pytest

from oracle_api import oracle_api
from topoptspline3d_amd import _abi
from topoptspline3d_amd.context import Context
from topoptspline3d_amd.domains import cantilever

E, NU, P, RMIN = 1e6, 0.4, 3.0, 1e-3


def spline_ctx(nx, ny, nz, coef_fn):
    solid, support, f = cantilever(nx, ny, nz)
    c = Context(oracle_api(), solid, support)
    c.set_material(E, NU, P, RMIN)
    n = c.spline_init(1, 4, 4)
    ncx, ncy, ncz = (nx + 3) // 4 + 1, (ny + 3) // 4 + 1, (nz + 3) // 4 + 1
    assert n == ncx * ncy * ncz
    k, j, i = np.meshgrid(np.arange(ncz), np.arange(ncy), np.arange(ncx), indexing="ij")
    c.set_coeffs(coef_fn(i, j, k).astype(np.float32).ravel())
    return c


def eval_fixed(ctx, **kw):
    o = _abi.SurfaceOpts()
    d = dict(cells=64, beta=8.0, print_angle=0.75 * np.pi, ss_mode=4, drip_mode=4, p_norm=8.0, hfunction_c=10.0,
             sigmoid_c=0.09, drip_beta=128.0, drip_mu=32.0, drip_angle=160 / 180 * np.pi, drip_alpha=4.0,
             drip_beta_scale=10.0, drip_version=0)
    d.update(kw)
    for key, v in d.items():
        setattr(o, key, v)
    ss, dr = C.c_double(), C.c_double()
    fn = oracle_api()._lib.orc_surface_eval_fixed
    fn.argtypes = [C.c_void_p, C.POINTER(_abi.SurfaceOpts), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    assert fn(ctx.h, C.byref(o), C.byref(ss), C.byref(dr)) == 0
    return ss.value, dr.value


def test_plane_level_set_hand_computed():
    """Coefficients 0.5 + 0.1 (k - 2) (linear in z) on a 16^3 box: the cubic spline
    reproduces the linear field, so {spline = 1/2} is the plane z = 8 (lattice index 2 of
    spacing 4).  An 8-cell background lattice (cell 2 voxels, nodes at z = 2k - 1) crosses it
    on the 8 x 8 vertical edges inside the box: 64 points at z = 8 exactly, plus the box-top
    crossings (nodes outside the box carry -1/2).  Solid lies above: stepping along grad
    spline enters solid, so the orientation is -1 and every normal is (0, 0, -|grad|); with
    the print angle 3 pi/4, inner = -1 / cos(3 pi/4) = sqrt 2 at every point, and the
    overhang mode's mean is 1 + sigmoid((sqrt 2 - 1.5) / 0.09)."""
    c = spline_ctx(16, 16, 16, lambda i, j, k: 0.5 + 0.1 * (k - 2.0))
    st = c.surface_constraints(cells=8, ss_mode=4, drip_mode=4)
    p, nv = c.surface_points(st["n_points"])
    plane = np.abs(p[:, 2] - 8.0) < 1.0
    assert plane.sum() == 64
    assert np.allclose(p[plane, 2], 8.0, atol=1e-12)
    assert sorted(set(np.round(p[plane, 0], 9))) == list(np.arange(1.0, 16.0, 2.0))
    assert st["direction"] == -1.0
    assert np.allclose(nv[:, :2], 0, atol=1e-12) and np.all(nv[:, 2] < 0)
    assert np.allclose(nv[plane, 2], -0.1 / 4, rtol=1e-12)
    assert st["n_down"] == st["n_points"]
    want = 1.0 + 1.0 / (1.0 + np.exp(-(np.sqrt(2.0) - 1.5) / 0.09))
    assert abs(st["ss_value"] - want) <= 1e-12
    # p-norm mode: (mean inner^8)^(1/8) - 1 = sqrt 2 - 1
    st0 = c.surface_constraints(cells=8, ss_mode=0, drip_mode=4)
    assert abs(st0["ss_value"] - (np.sqrt(2.0) - 1.0)) <= 1e-12
    # the reference's drip modes other than 6 and 7 give per-point value 0
    assert st["drip_value"] == 0.0
    assert c.surface_constraints(cells=8, drip_mode=0)["drip_value"] == -1.0


def wavy(i, j, k):
    # a perturbed ball: the level set 1/2 is a closed surface inside the 20x16x16 box
    return (0.95 - 0.12 * ((i - 2.5) ** 2 + (j - 2.0) ** 2 + (k - 2.0) ** 2)
            + 0.06 * np.sin(0.9 * i + 0.3) * np.cos(0.7 * j - 0.2) + 0.04 * np.cos(0.8 * k + 0.1 * i))


MODES = [dict(ss_mode=m, drip_mode=4) for m in range(6)] + [
    dict(ss_mode=4, drip_mode=6, drip_beta=4.0, drip_mu=2.0, drip_beta_scale=1.0),
    dict(ss_mode=5, drip_mode=7, drip_beta=4.0, drip_mu=2.0, drip_beta_scale=1.0),
    dict(ss_mode=4, drip_mode=7, drip_beta=4.0, drip_version=1, drip_alpha=0.5, drip_beta_scale=1.0),
    dict(ss_mode=3, drip_mode=7, drip_beta=4.0, drip_mu=2.0, drip_angle=0.3 * np.pi, drip_beta_scale=1.0),
]


@pytest.mark.parametrize("opts", MODES)
def test_gradients_match_finite_differences(opts):
    """d(value)/dc_j by central differences (step 2e-3 in FP32 coefficients) at the fixed
    points and orientation, for the coefficients with the largest gradients and a few
    random ones, every self-support mode and the drip modes (both orientations of the
    angle test and the Gauss variant)."""
    c = spline_ctx(20, 16, 16, wavy)
    kw = dict(cells=12, print_angle=0.75 * np.pi)
    kw.update(opts)
    st = c.surface_constraints(**kw)
    assert st["n_points"] > 50 and 0 < st["n_down"] < st["n_points"]
    coef = c.coeffs()
    for which, idx in ((2, 0), (3, 1)):
        if which == 3 and opts["drip_mode"] < 6:
            assert not np.any(c.coeff_sens(3))
            continue
        g = c.coeff_sens(which).astype(np.float64)
        assert np.any(g)
        rng = np.random.default_rng(which)
        js = list(np.argsort(-np.abs(g))[:6]) + list(rng.choice(np.nonzero(g)[0], 4))
        h = 2e-3
        for j in js:
            vals = []
            for s in (1, -1):
                cc = coef.copy()
                cc[j] += s * h
                c.set_coeffs(cc)
                vals.append(eval_fixed(c, **kw)[idx])
            c.set_coeffs(coef)
            # the perturbation is what FP32 actually stored
            hp = float(np.float32(coef[j] + h) - coef[j])
            hm = float(coef[j] - np.float32(coef[j] - h))
            fd = (vals[0] - vals[1]) / (hp + hm)
            assert abs(fd - g[j]) <= 2e-3 * np.abs(g).max() + 1e-3 * abs(g[j]), (which, j, fd, g[j])
