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


class ConvergeCriteria:
    """Stop test of the reference drivers (snippet::converge_criteria, utils/snippet.h:156-213):
    after each iteration the 4-point change of the objective,
    fch = |c_k + c_(k-1) - c_(k-2) - c_(k-3)| / (c_(k-2) + c_(k-3)), and the same change of
    every constraint value must be below the threshold (the constraints' below 1e-3) for the
    iteration to count as converged; a converged iteration adds one to a counter, any other
    takes one away (not below 0); the test passes once the counter exceeds `max_counter`.
    Values before the first iteration are NaN, so the first three iterations never count.
    The drivers construct it as (1, 5, 1e-3) (utils/test_utils.cpp:640, 1031)."""

    N = 20  # history length of the reference's circular arrays

    def __init__(self, n_constraints=1, max_counter=5, thres=1e-3):
        self.nc, self.max_counter, self.thres = n_constraints, max_counter, thres
        self.c = [float("nan")] * self.N
        self.g = [[float("nan")] * self.N for _ in range(n_constraints)]
        self.itn = 0
        self.counter = 0
        self.fch = float("nan")

    @staticmethod
    def _change(a, k):
        n = len(a)
        v = lambda j: a[j % n]
        return abs(v(k) + v(k - 1) - v(k - 2) - v(k - 3)) / (v(k - 2) + v(k - 3))

    def update(self, c, constraints):
        """record iteration values; returns True once the run counts as converged"""
        k = self.itn
        self.c[k % self.N] = c
        for i in range(self.nc):
            self.g[i][k % self.N] = constraints[i]
        self.fch = self._change(self.c, k)
        conv = self.fch < self.thres  # False for NaN
        if conv:
            for i in range(self.nc):
                if self._change(self.g[i], k) > 1e-3:  # NaN does not break the test here
                    conv = False
                    break
        self.counter = self.counter + 1 if conv else max(self.counter - 1, 0)
        self.itn += 1
        return self.counter > self.max_counter


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
    """With `heaviside` (the reference's ENABLE_HEAVISIDE build): the design variable is
    projected with beta = 8 before the first solve; each iteration takes the compliance
    and volume sensitivities through the projection's derivative before filtering
    (computeSensitivity2), updates the design variable by the projected OC
    (updateDensities2) and doubles beta, up to 64, when the mean design change falls
    below 1e-3 or every 5th iteration (utils/test_utils.cpp:680-684).  The record holds
    the reference's stop test (ConvergeCriteria(1, 5, 1e-3)) per iteration; configs fix
    the iteration count, so it is reported, not used."""
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
    rec = {"c": [], "v": [], "its": [], "rel_res": [], "ms": [], "g": [], "beta": [], "change": [],
           "stop": []}
    stop = ConvergeCriteria(1, 5, 1e-3)
    ph = Phases()
    for it in range(1, iters + 1):
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
        rec["stop"].append(stop.update(c, [v]))
        rec["beta"].append(beta)
        if heaviside:
            g, _, chmax, chmean = _timed(ph, "oc", ctx.oc_update_projected, p["vfrac"], p["move"], p["damp"], beta)
            rec["change"].append((chmax, chmean))
            if (chmean < 1e-3 or it % 5 == 0) and beta < 64:
                beta *= 2
        else:
            g, _ = _timed(ph, "oc", ctx.oc_update, p["vfrac"], p["move"], p["damp"])
        rec["c"].append(c); rec["v"].append(v); rec["g"].append(g)
        rec["its"].append(st["iterations"][0]); rec["rel_res"].append(st["rel_res"][0])
        rec["ms"].append((time.perf_counter() - t0) * 1e3)
    rec["phases_ms"] = {k: float(np.median(v)) for k, v in ph.t.items()}
    rec["ctx"] = ctx
    return rec


def run_spline_mma(api, nx, ny, nz, iters=100, rtol=1e-6, precision=1, spacing=4, order=4,
                   move=0.06, seed=11, maxit=200, sens_scale=None, vol_scale=1e3, heaviside=False,
                   self_support=False, surface=None, c0=None, **kw):
    """Cubic uniform spline lattice every `spacing` voxels, coefficients 0.4 +- 0.05
    (seeded), MMA on the coefficients with move limit `move`.  The constraint is scaled as
    in the reference driver (volScale 1e3, utils/test_utils.cpp:1015-1020); the objective
    is divided by its first value (sens_scale=None), because with a unit total load the
    raw compliance gradients sit below MMA's fixed 1e-5 regularisation.  With `heaviside`
    the spline density is projected (beta 4, doubled every 10th iteration up to 32,
    utils/test_utils.cpp:1014, 1076-1080, 1192-1196) and both sensitivities go through the
    projection's derivative before the filter.  The record holds the reference's stop
    test per iteration (reported, not used).  With `self_support` (the reference's
    ENABLE_SELFSUPPORT build, utils/test_utils.cpp:1034-1188): MMA with m = 3, the
    self-supporting and drip constraints evaluated on the spline's surface points each
    iteration (`surface`: keyword options of Context.surface_constraints), scaled by
    SSScale / dripScale = 1e-1 / 1e2, from iteration 31 1e0 / 1e3 and from iteration 81
    1e1 / 1e4.  While all coefficients are equal the constraints read ss = 1, drip = 0
    with zero gradients (grids.areALLCoeffsEqual, utils/test_utils.cpp:1106-1113).
    `c0`: initial coefficients in place of the seeded ones."""
    p = dict(DEFAULTS, **kw)
    solid, support, f = cantilever(nx, ny, nz)
    ctx = Context(api, solid, support)
    ctx.set_material(p["E"], p["nu"], p["penal"], p["rho_min"])
    ctx.set_loads(f)
    n = ctx.spline_init(1, order, spacing)
    rng = np.random.default_rng(seed)
    if c0 is None:
        c0 = (p["vfrac"] + rng.uniform(-0.05, 0.05, n)).astype(np.float32)
    c0 = np.asarray(c0, np.float32)
    ctx.set_coeffs(c0)
    mma = MMA(api, n, 3 if self_support else 1, p["rho_min"], 1.0, move=move)
    mma.set_x(c0.astype(np.float64))
    rec = {"c": [], "v": [], "its": [], "rel_res": [], "ms": [], "beta": [], "stop": [], "ss": [], "drip": []}
    stop = ConvergeCriteria(3 if self_support else 1, 5, 1e-3)
    beta = 4.0
    ph = Phases()
    for it in range(1, iters + 1):
        t0 = time.perf_counter()
        ctx.spline_eval()
        if heaviside:
            ctx.project_density(beta)
        v = ctx.volume()
        _timed(ph, "operator", ctx.update_operator)
        st = _timed(ph, "solve", ctx.solve, rtol=rtol, maxit=maxit, precision=precision)
        c = ctx.compliance()[0]
        t1 = time.perf_counter()
        ss_val, drip_val = 1.0, 0.0
        if self_support:
            ss_scale, drip_scale = (1e1, 1e4) if it > 80 else (1e0, 1e3) if it > 30 else (1e-1, 1e2)
            st_s = _timed(ph, "surface", ctx.surface_constraints, beta=beta, **(surface or {}))
            ss_val, drip_val = st_s["ss_value"], st_s["drip_value"]
            rec["ss"].append(ss_val); rec["drip"].append(drip_val)
        rec["stop"].append(stop.update(c, [v, ss_val, drip_val] if self_support else [v]))
        rec["beta"].append(beta)
        ctx.sensitivity()
        ctx.init_vol_sensitivity()
        if heaviside:
            ctx.projection_sensitivity(beta, 0)
            ctx.projection_sensitivity(beta, 1)
        ctx.filter(p["radius"], 0)
        ctx.filter(p["radius"], 1)
        ctx.spline_adjoint(0)
        ctx.spline_adjoint(1)
        ph.add("sens_filter_adjoint", time.perf_counter() - t1)
        if sens_scale is None:
            sens_scale = 1.0 / c
        if self_support:
            _timed(ph, "mma", mma.step_on_surface, ctx, sens_scale, vol_scale, p["vfrac"], ss_scale, drip_scale)
        else:
            _timed(ph, "mma", mma.step_on, ctx, sens_scale, vol_scale, p["vfrac"])
        if heaviside and it % 10 == 0 and it > 2 and beta < 32:
            beta *= 2
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
