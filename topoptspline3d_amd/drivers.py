"""Optimisation loops over the C ABI, in the order of the reference drivers.

run_oc       -- TestSuit::testOrdinaryTopopt (utils/test_utils.cpp:604-706): per
                iteration update_stencil, solve, compliance/volume, sensitivity, filter,
                OC update (Heaviside projection optional, off by default).
run_spline_mma -- TestSuit::testOrdinarySplineTopoptMMA (utils/test_utils.cpp:972-1217):
                coefficients -> density, solve, sensitivity, filter, spline adjoint of the
                compliance and volume sensitivities, MMA step on the coefficients.
run_worst_oc -- optimization() (optimization/optimization.cpp:673-767), the reference's
                default loop: per iteration update_stencil, worst-case load by the modified
                power method over a load region, its sensitivity, filter, OC update.
All return per-iteration records (compliance, volume, solver iterations, ms per phase).
"""
import os
import time

import numpy as np

from . import bio
from .context import MMA, Context
from .domains import cantilever, face_load_region

DEFAULTS = dict(E=1e6, nu=0.4, penal=3.0, rho_min=1e-3, vfrac=0.4, radius=2.0, move=0.03, damp=0.5)


class Phases:
    def __init__(self):
        self.t = {}

    def add(self, name, dt):
        self.t.setdefault(name, []).append(dt * 1e3)


def _timed(ph, name, fn, *a, **k):
    t0 = time.perf_counter()
    r = fn(*a, **k)
    ph.add(name, time.perf_counter() - t0)
    return r


def run_oc(api, nx, ny, nz, iters=50, rtol=1e-8, precision=0, maxit=200, domain=None,
           heaviside=False, **kw):
    p = dict(DEFAULTS, **kw)
    if domain is None:
        solid, support, f = cantilever(nx, ny, nz)
        shell = None
    else:
        solid, support, f, shell = domain
    ctx = Context(api, solid, support, shell=shell)
    ctx.set_material(p["E"], p["nu"], p["penal"], p["rho_min"])
    ctx.set_loads(f)
    ctx.set_density(np.where(solid.ravel() > 0, p["vfrac"], 0).astype(np.float32))
    beta = 8.0
    if heaviside:
        ctx.project_density(beta)
    rec = {"c": [], "v": [], "its": [], "rel_res": [], "ms": [], "g": []}
    ph = Phases()
    for it in range(iters):
        t0 = time.perf_counter()
        _timed(ph, "operator", ctx.update_operator)
        st = _timed(ph, "solve", ctx.solve, rtol=rtol, maxit=maxit, precision=precision)
        c = _timed(ph, "compliance", ctx.compliance)[0]
        v = ctx.volume()
        t1 = time.perf_counter()
        ctx.sensitivity()
        if heaviside:
            ctx.init_vol_sensitivity()
            ctx.projection_sensitivity(beta, 0)
            ctx.projection_sensitivity(beta, 1)
        ctx.filter(p["radius"], 0)
        if heaviside:
            ctx.filter(p["radius"], 1)
        ph.add("sens_filter", time.perf_counter() - t1)
        g, _ = _timed(ph, "oc", ctx.oc_update, p["vfrac"], p["move"], p["damp"])
        if heaviside:
            ctx.project_density(beta)
            if (it + 1) % 5 == 0 and beta < 64:
                beta *= 2
        rec["c"].append(c); rec["v"].append(v); rec["g"].append(g)
        rec["its"].append(st["iterations"][0]); rec["rel_res"].append(st["rel_res"][0])
        rec["ms"].append((time.perf_counter() - t0) * 1e3)
    rec["phases_ms"] = {k: float(np.median(v)) for k, v in ph.t.items()}
    rec["ctx"] = ctx
    return rec


def run_spline_mma(api, nx, ny, nz, iters=100, rtol=1e-6, precision=1, spacing=4, order=4,
                   move=0.06, seed=11, maxit=200, sens_scale=None, vol_scale=1e3, **kw):
    """Cubic uniform spline lattice every `spacing` voxels, coefficients 0.4 +- 0.05
    (seeded), MMA on the coefficients with move limit `move`.  The constraint is scaled as
    in the reference driver (volScale 1e3, utils/test_utils.cpp:1015-1020); the objective
    is divided by its first value (sens_scale=None), because with a unit total load the
    raw compliance gradients sit below MMA's fixed 1e-5 regularisation."""
    p = dict(DEFAULTS, **kw)
    solid, support, f = cantilever(nx, ny, nz)
    ctx = Context(api, solid, support)
    ctx.set_material(p["E"], p["nu"], p["penal"], p["rho_min"])
    ctx.set_loads(f)
    n = ctx.spline_init(1, order, spacing)
    rng = np.random.default_rng(seed)
    c0 = (p["vfrac"] + rng.uniform(-0.05, 0.05, n)).astype(np.float32)
    ctx.set_coeffs(c0)
    mma = MMA(api, n, 1, p["rho_min"], 1.0, move=move)
    mma.set_x(c0.astype(np.float64))
    rec = {"c": [], "v": [], "its": [], "rel_res": [], "ms": []}
    ph = Phases()
    for it in range(iters):
        t0 = time.perf_counter()
        ctx.spline_eval()
        v = ctx.volume()
        _timed(ph, "operator", ctx.update_operator)
        st = _timed(ph, "solve", ctx.solve, rtol=rtol, maxit=maxit, precision=precision)
        c = ctx.compliance()[0]
        t1 = time.perf_counter()
        ctx.sensitivity()
        ctx.init_vol_sensitivity()
        ctx.filter(p["radius"], 0)
        ctx.filter(p["radius"], 1)
        ctx.spline_adjoint(0)
        ctx.spline_adjoint(1)
        ph.add("sens_filter_adjoint", time.perf_counter() - t1)
        if sens_scale is None:
            sens_scale = 1.0 / c
        _timed(ph, "mma", mma.step_on, ctx, sens_scale, vol_scale, p["vfrac"])
        rec["c"].append(c); rec["v"].append(v)
        rec["its"].append(st["iterations"][0]); rec["rel_res"].append(st["rel_res"][0])
        rec["ms"].append((time.perf_counter() - t0) * 1e3)
    rec["phases_ms"] = {k: float(np.median(v)) for k, v in ph.t.items()}
    rec["ctx"] = ctx
    return rec


def run_worst_oc(api, nx, ny, nz, iters=20, face="x+", inset=0, free_direction=False, precision=0,
                 out_dir=None, domain=None, **kw):
    """Box with x = 0 fixed; load region: the nodes of one face (`face_load_region`), force
    along the face normal unless `free_direction`.  Iteration i (from 1) draws its start
    force with seed i (the reference draws a new random force each time, Grid::randForce).
    With `out_dir`, writes the reference's records there (optimization.cpp:728-765):
    iter<i>_fs and crec_iter / vrec_iter / trec_iter every iteration, then cworst, vrec,
    trec, flast, ulast.  `domain`: a dict of `mesh_domain.load_config` (solid, support,
    shell, load_nodes, normals) in place of the box (nx, ny, nz are then ignored)."""
    p = dict(DEFAULTS, **kw)
    shell = None
    if domain is None:
        solid, support, f = cantilever(nx, ny, nz)
        ids, nrm = face_load_region(nx, ny, nz, face, inset)
    else:
        solid, support, f, shell = domain["solid"], domain["support"], domain["f"], domain["shell"]
        ids, nrm = domain["load_nodes"], domain["normals"]
    ctx = Context(api, solid, support, shell=shell)
    ctx.set_material(p["E"], p["nu"], p["penal"], p["rho_min"])
    ctx.set_loads(np.zeros_like(f))
    ctx.set_density(np.where(solid.ravel() > 0, p["vfrac"], 0).astype(np.float32))
    rec = {"c": [], "v": [], "its": [], "ms": []}
    ph = Phases()
    for it in range(1, iters + 1):
        t0 = time.perf_counter()
        _timed(ph, "operator", ctx.update_operator)
        st = _timed(ph, "worst_case", ctx.worst_compliance, ids, None if free_direction else nrm, seed=it,
                    precision=precision)
        v = ctx.volume()
        t1 = time.perf_counter()
        ctx.sensitivity()
        ctx.filter(p["radius"], 0)
        ph.add("sens_filter", time.perf_counter() - t1)
        _timed(ph, "oc", ctx.oc_update, p["vfrac"], p["move"], p["damp"])
        rec["c"].append(st["compliance"]); rec["v"].append(v); rec["its"].append(st["iterations"])
        rec["ms"].append((time.perf_counter() - t0) * 1e3)
        if out_dir:
            bio.write_support_force(os.path.join(out_dir, f"iter{it}_fs"), ctx, ids)
            bio.write_history(out_dir, rec)
    if out_dir:
        bio.write_history(out_dir, rec, final=True)
        bio.write_support_force(os.path.join(out_dir, "flast"), ctx, ids)
        bio.write_displacement(os.path.join(out_dir, "ulast"), ctx)
    rec["phases_ms"] = {k: float(np.median(v)) for k, v in ph.t.items()}
    rec["ctx"] = ctx
    rec["load_nodes"] = ids
    return rec
