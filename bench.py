"""Benchmark: ms per topology-optimisation iteration on one MI355X.

Default workload (the production-size case the headline metric is quoted on): the
irregular 511x320x320 ellipsoid-union domain (~36% solid, z = 0 supported, unit load on
the top surface along the occupancy normal, 3-voxel shell), cubic B-spline density on a
lattice every 4 voxels, MMA on the coefficients, 7-level geometric MG.  One step = one
optimisation iteration: spline evaluation, operator update (rho^p, Galerkin levels,
coarsest factor), flexible PCG to 1e-6 relative residual, compliance, sensitivity,
filter, spline adjoint of the compliance and volume sensitivities, MMA step.  All inputs
are resident on the GPU before the timed region.

Three runs start from the same initial design and cover the same iterations (`warmup`
untimed, then `steps` timed):
  value    -- plain mixed-precision MG-PCG (FP32 V-cycle reading the coarse levels'
              stencils as FP16 -- precision 2 --, FP64 outer residual);
  fp64     -- the same optimisation with FP64 V-cycles (plain FP64 GPU multigrid);
  learned  -- as `value` with the learned level-2 correction attached (trained before
              the timed region on pairs logged from 200 optimisation iterations on 128^3
              ellipsoid domains).
Then, on the final operator of the first run: solves from zero in FP64, mixed FP32 and
FP16 stencils and learned mode; the worst-case load search (modified power method); the
16-load-case batched workload; and the CPU baseline, one complete optimisation iteration
of the FP64 oracle on the host cores (including the MMA step).

Prints one JSON line.  Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
[--workload c2|c1|c0|fertility] [--no-cpu-baseline] [--no-learned] [--no-fp64-compare]
[--learn-only]
(fertility: the reference's bench mesh fertility/config1, kept under tests/golden,
voxelised at resolution 511 with its JSON supports and loads, element-density OC)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ms/TO-iter at 511-reso to 1e-6 residual, learned vs plain MG; HBM GB/s"
HBM_PEAK = 8000.0  # GB/s, MI355X
# TFLOP/s, MI355X vector FMA: FP64 78.6 (published); FP32 157.3 published with packed FMAs,
# 78.6 without them (packed FP32 is compiled out)
FP32_VECTOR_PEAK, FP64_VECTOR_PEAK = 78.6, 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def active_nodes(solid):
    nz, ny, nx = solid.shape
    act = np.zeros((nz + 1, ny + 1, nx + 1), bool)
    s = solid.astype(bool)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                act[dz:dz + nz, dy:dy + ny, dx:dx + nx] |= s
    return int(act.sum())


def make_workload(name):
    from topoptspline3d_amd.domains import cantilever, ellipsoids
    if name == "c2":
        solid, support, f, shell, meta = ellipsoids(511, 320, 320, seed=2024)
        return dict(name="c2", desc="ellipsoid union 511x320x320, spline spacing 4 (cubic), MMA",
                    solid=solid, support=support, f=f, shell=shell, meta=meta, spline=(1, 4, 4),
                    optimizer="mma", move=0.06, rtol=1e-6)
    if name == "c2slab":
        # one GPU's share of an 8-way x-slab split of the ellipsoid workload: 64 of the 511
        # element columns, taken where the body stands on its supports (z = 0 touches the
        # solid for x in 331..464) so that the slab alone is a supported problem, keeping
        # the part connected to those supports inside the slab; same spline / MMA loop
        from scipy import ndimage
        solid, support, f, shell, meta = ellipsoids(511, 320, 320, seed=2024)
        x0, w = 368, 64
        sol = solid[:, :, x0:x0 + w].astype(bool)
        lab, _ = ndimage.label(sol)
        keep = np.zeros(lab.max() + 1, bool)
        keep[np.unique(lab[0])] = True
        keep[0] = False
        sol = keep[lab]
        nz, ny = sol.shape[:2]
        act = np.zeros((nz + 1, ny + 1, w + 1), bool)
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    act[dz:dz + nz, dy:dy + ny, dx:dx + w] |= sol
        sup = support.reshape(nz + 1, ny + 1, -1)[:, :, x0:x0 + w + 1] & act
        ff = f.reshape(3, nz + 1, ny + 1, -1)[:, :, :, x0:x0 + w + 1] * act
        sh = shell[:, :, x0:x0 + w] & sol
        return dict(name="c2slab", desc="x-slab 64x320x320 (columns 368..431) of the ellipsoid union 511x320x320, "
                    "spline spacing 4 (cubic), MMA", solid=sol.astype(np.uint8), support=sup.astype(np.uint8),
                    f=ff.reshape(-1), shell=sh.astype(np.uint8), meta={}, spline=(1, 4, 4), optimizer="mma",
                    move=0.06, rtol=1e-6)
    if name == "c1":
        solid, support, f = cantilever(128, 64, 64)
        return dict(name="c1", desc="cantilever 128x64x64, spline spacing 4 (cubic), MMA",
                    solid=solid, support=support, f=f, shell=None, meta={}, spline=(1, 4, 4),
                    optimizer="mma", move=0.06, rtol=1e-6)
    if name == "fertility":
        # a reference bench input at the headline resolution: mesh + JSON regions and loads
        from topoptspline3d_amd.mesh_domain import load_config
        d = load_config(os.path.join(ROOT, "tests", "golden", "bench_fertility", "config1.json"), 511,
                        shell_width=3)
        return dict(name="fertility", desc="bench mesh fertility (config1) voxelised at resolution 511, "
                    "3-voxel shell, element densities, OC", solid=d["solid"], support=d["support"], f=d["f"],
                    shell=d["shell"], meta={}, spline=None, optimizer="oc", move=0.03, rtol=1e-6)
    if name == "c0":
        solid, support, f = cantilever(24, 12, 12)
        return dict(name="c0", desc="cantilever 24x12x12, element densities, OC", solid=solid,
                    support=support, f=f, shell=None, meta={}, spline=None, optimizer="oc",
                    move=0.03, rtol=1e-8)
    raise SystemExit(f"unknown workload {name}")


class Loop:
    """One optimisation iteration over a context (api = HIP or oracle binding)."""

    def __init__(self, api, wl, precision, recycle=0):
        from topoptspline3d_amd.context import MMA, Context
        self.api, self.wl, self.precision = api, wl, precision
        ctx = Context(api, wl["solid"], wl["support"], shell=wl["shell"])
        ctx.set_material(1e6, 0.4, 3.0, 1e-3)
        ctx.set_loads(wl["f"])
        ctx.set_start_recycling(recycle)  # 0: plain warm start
        self.ctx = ctx
        if wl["spline"]:
            mode, order, spacing = wl["spline"]
            n = ctx.spline_init(mode, order, spacing)
            rng = np.random.default_rng(11)
            c0 = (0.4 + rng.uniform(-0.05, 0.05, n)).astype(np.float32)
            ctx.set_coeffs(c0)
            self.mma = MMA(api, n, 1, 1e-3, 1.0, move=wl["move"])
            self.mma.set_x(c0.astype(np.float64))
        else:
            ctx.set_density(np.where(wl["solid"].ravel() > 0, 0.4, 0).astype(np.float32))
        self.its, self.res, self.c = [], [], []
        self.sens_scale = None

    def step(self, phases=None):
        """one optimisation iteration; `phases` (a dict) receives seconds per phase"""
        ctx, wl = self.ctx, self.wl
        t = [time.perf_counter()]

        def mark(name):
            t.append(time.perf_counter())
            if phases is not None:
                phases[name] = phases.get(name, 0.0) + t[-1] - t[-2]

        if wl["spline"]:
            ctx.spline_eval()
            mark("spline eval")
        # FP32 cycles: coarse operators built into FP32 storage; FP64 cycles: FP64
        ctx.set_galerkin_precision(1 if self.precision else 0)
        ctx.update_operator()
        mark("operator update")
        st = ctx.solve(rtol=wl["rtol"], maxit=300, precision=self.precision)
        self.its.append(st["iterations"][0])
        self.res.append(st["rel_res"][0])
        self.c.append(ctx.compliance()[0])
        mark("solve")
        ctx.sensitivity()
        ctx.filter(2.0, 0)
        if wl["spline"]:
            ctx.init_vol_sensitivity()
            ctx.filter(2.0, 1)
            ctx.spline_adjoint(0)
            ctx.spline_adjoint(1)
            mark("sensitivity/filter/adjoint")
            if self.sens_scale is None:  # objective normalised by its first value
                self.sens_scale = 1.0 / self.c[-1]
            self.mma.step_on(ctx, self.sens_scale, 1e3, 0.4)
            mark("MMA")
        else:
            mark("sensitivity/filter")
            ctx.oc_update(0.4, wl["move"], 0.5)
            mark("OC")


def sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def kernel_stats(api, ctx, enable):
    import ctypes
    from topoptspline3d_amd import _abi
    ks = _abi.KernelStats()
    rc = api.profile(ctx.h, 1 if enable else 0, ctypes.byref(ks))
    if rc:
        raise RuntimeError("profile failed")
    if enable:
        return None
    return {ks.names[i].value.decode(): (ks.total_ms[i], ks.launches[i]) for i in range(ks.n_classes)}


DTYPES = {0: "f64", 1: "f32 V-cycle / f64 PCG", 2: "f32 V-cycle (f16 coarse stencils) / f64 PCG",
          3: "f32 V-cycle (f16 coarse stencils) / f32 PCG residual, f64 u and operator"}
SOLVERS = {0: "FP64 V-cycle", 1: "FP32 V-cycle", 2: "FP32 V-cycle on FP16 coarse stencils",
           3: "FP32 V-cycle on FP16 coarse stencils, FP32 residual and K p"}


def roofline(stats, info, n_active_nodes, steps, workload_tag=None, precision=1, sweep_mode=0):
    """Roofline of the dominant kernel class (by total time among the classes with a
    single-level definition).  Algorithmic bytes per launch (each array read or written
    once at storage precision, index/flag re-reads excluded):
      fine GS (FP32): one sweep = 4 N_e + (9*4 + 4) N_v; one launch = 1/8 (a colour) or,
        with the in-wave paired schedule (sweep mode 3), 1/4 (two colours)
      fine residual (FP32/FP64):  4 N_e + (9 s + 4) N_v
      fine operator (FP64):       4 N_e + (6*8 + 4) N_v
      level-1 stencil GS colour pass (FP32 cycle): (243*4 + 9*4) N_c1 / 8, or with the
        FP16 stencil copy (precision 2) (122*4 + 4 + 9*4) N_c1 / 8"""
    Ne, Nv = info["n_active_elements"][0], n_active_nodes
    Nc1 = info["n_free_nodes"][1] if info["n_levels"] > 1 else 0
    colours = 2 if sweep_mode == 3 else 1  # colours per fine FP32 GS launch
    per_launch = {
        "fine_gs_f32": (4 * Ne + (9 * 4 + 4) * Nv) / 8.0 * colours,
        "fine_gs_f64": (4 * Ne + (9 * 8 + 4) * Nv) / 8.0,
        "fine_resid_f32": 4 * Ne + (9 * 4 + 4) * Nv,
        "fine_resid_f64": 4 * Ne + (9 * 8 + 4) * Nv,
        "fine_apply_f64": 4 * Ne + (6 * 8 + 4) * Nv,
        "st_gs_f32_l1": ((122 * 4 + 4 if precision >= 2 else 243 * 4) + 9 * 4) * Nc1 / 8.0,
    }
    flops = {"fine_gs_f32": 1170 * Nv / 8.0 * colours, "fine_gs_f64": 1170 * Nv / 8.0, "fine_resid_f32": 1152 * Nv,
             "fine_resid_f64": 1152 * Nv, "fine_apply_f64": 1152 * Nv, "st_gs_f32_l1": 504 * Nc1 / 8.0}
    dom = max((k for k in stats if k in per_launch), key=lambda k: stats[k][0])
    tot = {k: v[0] for k, v in stats.items() if v[1] > 0}
    out = {"bound": "hbm", "peak": HBM_PEAK, "unit": "GB/s", "traffic": None, "kernel": dom,
           "kernel_ms_per_step": {k: round(v / steps, 3) for k, v in tot.items()}}
    # HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE counters (scaled
    # by a device-copy calibration) for this kernel class and workload, when recorded
    tf = os.path.join(ROOT, "profiles", f"{dom}_{workload_tag}.json") if workload_tag else None
    if tf and os.path.exists(tf):
        js = json.load(open(tf))
        t = js.get("traffic")
        # (a GS profile holds for the schedule it was taken with)
        if t and (not dom.startswith("fine_gs") or int(js.get("sweep_mode", 0)) == (sweep_mode if dom == "fine_gs_f32" else 0)):
            out["traffic"] = round(t["bytes_per_launch"])
            out["traffic_source"] = os.path.relpath(tf, ROOT)
            if js.get("valu"):
                out["valu_busy_counter"] = js["valu"]["valu_busy"]
    if dom in per_launch and stats[dom][1] > 0:
        avg_ms = stats[dom][0] / stats[dom][1]
        ach = per_launch[dom] / (avg_ms * 1e-3) / 1e9
        out.update({"achieved": round(ach, 1), "frac": round(ach / HBM_PEAK, 4),
                    "avg_launch_us": round(avg_ms * 1e3, 2), "launches": stats[dom][1],
                    "bytes_per_launch": per_launch[dom],
                    "achieved_tflops": round(flops[dom] / (avg_ms * 1e-3) / 1e12, 2)})
        # the same launches against the vector FP32/FP64 rate (no packed FP32 FMAs)
        vpeak = FP32_VECTOR_PEAK if dom.endswith("f32") or dom == "st_gs_f32_l1" else FP64_VECTOR_PEAK
        out["compute"] = {"achieved_tflops": out["achieved_tflops"], "peak_tflops": vpeak,
                          "frac": round(out["achieved_tflops"] / vpeak, 4),
                          "peak_note": "vector FMA rate without packed FP32"}
    else:
        out.update({"achieved": None, "frac": None})
    return out


def cpu_baseline(wl, gpu_its, coeffs=None):
    """FP64 oracle on the host cores: one complete optimisation iteration of the same
    workload from the same initial design (spline evaluation, operator update, FCG solve
    from zero to the workload's tolerance with FP64 V-cycles, compliance, sensitivity,
    filter, spline adjoint and MMA step -- or OC for element-density workloads), timed
    per phase.  With `coeffs` (the GPU run's final coefficients), the surface constraints
    on the 64-cell lattice are timed on them as well (reported beside, not in the value)."""
    import ctypes
    from topoptspline3d_amd import _abi
    lib = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(lib):
        import subprocess
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    api = _abi.bind(ctypes.CDLL(lib), "orc_")
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    loop = Loop(api, wl, 0)
    t_setup = time.perf_counter() - t0
    ph = {}
    t_iter = time.perf_counter()
    loop.step(phases=ph)
    t_iter = time.perf_counter() - t_iter
    its = loop.its[-1]
    surf = ""
    if coeffs is not None:
        loop.ctx.set_coeffs(coeffs)
        t = time.perf_counter()
        st = loop.ctx.surface_constraints(cells=64, beta=8.0)
        surf = (f"; surface constraints on the GPU run's final coefficients, 64 cells: "
                f"{(time.perf_counter() - t) * 1e3:.1f} ms, {st['n_points']} points")
    loop.ctx.close()
    cpu = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    parts = ", ".join(f"{k} {v:.2f}s" for k, v in ph.items())
    return {"value": round(t_iter * 1e3, 1), "unit": "ms/TO-iter", "cores": cores, "kind": "port",
            "sample": (f"one complete optimisation iteration of the FP64 oracle on the same domain from the same "
                       f"initial design ({parts}; FCG from zero: {its} iterations, the GPU's timed iterations "
                       f"took {min(gpu_its)}-{max(gpu_its)} from their warm starts); context set-up "
                       f"{t_setup:.1f}s not included{surf}; CPU: {cpu or 'unknown'}")}


def learned_phase(api, n_dom, iters, steps, level=2):
    """Training of the learned level-2 correction: residual / error pairs logged from the
    solver's own V-cycles over `iters` spline + MMA iterations on each of `n_dom` 128^3
    ellipsoid-union domains (semi-axes scaled to 15-45 voxels), then Adam on 90 % of the
    pairs; the rest is held out."""
    from topoptspline3d_amd import learned as LN
    from topoptspline3d_amd.domains import ellipsoids
    t0 = time.perf_counter()
    ds = None
    its = []
    for d in range(n_dom):
        solid, support, f, shell, meta = ellipsoids(128, 128, 128, seed=100 + d, axes=(15, 45))
        wl = dict(name="train", desc="", solid=solid, support=support, f=f, shell=shell, meta=meta,
                  spline=(1, 4, 4), optimizer="mma", move=0.06, rtol=1e-6)
        loop = Loop(api, wl, 1)
        if ds is None:
            ds = LN.PairDataset(api, LN.level_box(loop.ctx, level), n_dom * iters * 40)
        LN.log_pairs(loop.ctx, ds, level=level, n_inner=8, stride=1)
        for _ in range(iters):
            loop.step()
        its += loop.its
        LN.log_pairs(loop.ctx, None)
        loop.ctx.close()
        log(f"[bench] learned: domain {d} fill {meta.get('fill', 0):.3f}, pairs so far {len(ds)}")
    t_gen = time.perf_counter() - t0
    n = len(ds)
    ntr = max(1, int(0.9 * n))
    net = LN.UNet(api, seed=7)
    before = net.evaluate(ds, ntr, n - ntr) if n > ntr else None
    t = time.perf_counter()
    lf, ll = net.train(ds, 0, ntr, steps=steps, batch=8, lr=1e-3, seed=3)
    t_train = time.perf_counter() - t
    after = net.evaluate(ds, ntr, n - ntr) if n > ntr else None
    # training throughput against the BF16 dense MFMA peak (forward + data and weight
    # gradients ~ 3 forward passes per sample)
    train_flops = 3.0 * LN.unet_flops(ds.box) * steps * 8
    info = {"pairs": n, "box": list(ds.box), "domains": n_dom, "to_iterations": n_dom * iters,
            "mean_pcg_its": round(float(np.mean(its)), 2), "generate_s": round(t_gen, 1),
            "train_steps": steps, "batch": 8, "train_s": round(t_train, 2),
            "train_tflops": round(train_flops / t_train / 1e12, 2),
            "train_frac_of_bf16_peak": round(train_flops / t_train / 2.5e15, 4),
            "train_loss_first_last": [round(lf, 4), round(ll, 4)],
            "heldout_loss_untrained_trained": [round(before, 4) if before is not None else None,
                                               round(after, 4) if after is not None else None]}
    log(f"[bench] learned: {info}")
    return net, ds, info


def batched_phase(api, steps):
    """Batched worst-case loads: 256x128x128 cantilever, 16 unit patch loads of seeded
    directions on the x = max face, all cases in one interleaved FCG (graph-replayed
    iteration, autotuned block sizes); objective max_k c_k with the sensitivity of the
    argmax case, filter, OC.  Timed per optimisation iteration, plus the same 16 cases
    solved one at a time on the final operator."""
    from topoptspline3d_amd.context import Context
    from topoptspline3d_amd.domains import batched_patch_loads, cantilever
    nx, ny, nz, K = 256, 128, 128, 16
    solid, support, _ = cantilever(nx, ny, nz)
    F, _ = batched_patch_loads(nx, ny, nz, K=K, patch=8, seed=7)
    ctx = Context(api, solid, support)
    ctx.set_material(1e6, 0.4, 3.0, 1e-3)
    ctx.set_loads(F)
    ctx.set_density(np.full(solid.size, 0.4, np.float32))
    ctx.set_galerkin_precision(1)  # FP32 cycles only

    def it(i):
        ctx.update_operator()
        st = ctx.solve_batched(rtol=1e-6, maxit=300)
        c = ctx.compliance()
        k = int(np.argmax(c))
        ctx.sensitivity(k)
        ctx.filter(2.0, 0)
        ctx.oc_update(0.4, 0.03, 0.5)
        log(f"[bench] batched step {i}: max c {c[k]:.4e} (case {k}), its {min(st['iterations'])}-{max(st['iterations'])}")
        return st

    it(-1)  # warm-up: tuning and graph capture
    sync()
    t = time.perf_counter()
    sts = [it(i) for i in range(steps)]
    sync()
    ms = (time.perf_counter() - t) * 1e3 / steps
    ctx.update_operator()
    sb = ctx.solve_batched(rtol=1e-6, maxit=300)
    ss = ctx.solve(rtol=1e-6, maxit=300, precision=1, warm_start=False)
    out = {"workload": f"cantilever {nx}x{ny}x{nz}, K={K} patch loads, max-compliance OC",
           "ms_per_iter": round(ms, 2), "solve_ms_batched": round(sb["seconds"] * 1e3, 2),
           "solve_ms_one_case_at_a_time": round(ss["seconds"] * 1e3, 2),
           "batched_speedup": round(ss["seconds"] / max(sb["seconds"], 1e-9), 2),
           "iterations_batched": sb["iterations"], "iterations_one_at_a_time": ss["iterations"],
           "max_rel_res": max(sb["rel_res"]), "tuned_block_sizes": ctx.batch_tuning()}
    ctx.close()
    return out


def worst_phase(ctx, wl, precision):
    """Worst-case load over the workload's load region by the modified power method
    (modifiedPM, optimization/optimization.cpp:213-351): the loaded nodes, each force kept
    along that node's load vector, seeded start force, the reference's stop rule (force
    change 1e-4, cycle residual 1e-2, at most 500 V-cycles); on the final operator, with
    the mixed-precision cycles and with FP64 cycles.  Load case 0 is restored afterwards."""
    f = wl["f"].reshape(3, -1)
    nodes = np.nonzero(np.abs(f).sum(0) > 0)[0]
    normals = np.ascontiguousarray(f[:, nodes].T)
    out = {"load_nodes": int(nodes.size)}
    for tag, prec in (("mixed", precision), ("fp64", 0)):
        st = ctx.worst_compliance(nodes, normals=normals, seed=1, precision=prec)
        its = st["iterations"]
        out[tag] = {"cycles": its, "seconds": round(st["seconds"], 3),
                    "ms_per_cycle": round(st["seconds"] * 1e3 / max(its, 1), 3),
                    "worst_compliance": st["compliance"], "fch": st["fch"], "rel_res": st["rel_res"]}
        log(f"[bench] worst-case load ({tag}): {its} cycles, {st['seconds']:.2f} s, c {st['compliance']:.6e}")
    ctx.set_loads(wl["f"])
    return out


def surface_phase(ctx, reps=5, cells_list=(64, 128, 256)):
    """Self-supporting and drip constraints on the spline's surface points (deal_surface_points,
    optimization/optimization.cpp:1103-1128) for the current coefficients: the reference's
    64-cell background lattice and two finer ones, modes "oh"/"oh", beta 8; median of
    `reps` calls each (every call runs the whole chain: lattice values, vertex count and
    scan, points, orientation vote, constraint values and coefficient gradients)."""
    out = {}
    for cells in cells_list:
        ts = []
        for _ in range(reps):
            sync()
            t = time.perf_counter()
            st = ctx.surface_constraints(cells=cells, beta=8.0)
            sync()
            ts.append(time.perf_counter() - t)
        out[f"cells_{cells}"] = {"ms": round(float(np.median(ts)) * 1e3, 3), "points": int(st["n_points"]),
                                 "ss": st["ss_value"], "drip": st["drip_value"]}
        log(f"[bench] surface constraints, {cells} cells: {out[f'cells_{cells}']}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--precision", type=int, default=3,
                    help="0 FP64 V-cycle; 1 FP32 V-cycle; 2 FP32 V-cycle on FP16 coarse stencils; "
                         "3 as 2 with the FCG residual and K p stored in FP32")
    ap.add_argument("--recycle", type=int, default=4,
                    help="start of each solve projected onto the warm start and the earlier solutions "
                         "(basis size; 0: plain warm start)")
    ap.add_argument("--no-recycle-compare", action="store_true",
                    help="skip the run from the plain warm start")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64-compare", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-worst", action="store_true", help="skip the worst-case load search")
    ap.add_argument("--no-surface", action="store_true", help="skip the surface-constraint timing")
    ap.add_argument("--probe-element-mfma", action="store_true",
                    help="only the warmup iterations and the MFMA element-product measurement")
    ap.add_argument("--surface-only", action="store_true",
                    help="only the warmup iterations and the surface-constraint timing (profiling)")
    ap.add_argument("--surface-cells", type=int, nargs="*", default=[64, 128, 256],
                    help="background lattice sizes of the surface-constraint timing")
    ap.add_argument("--learn-domains", type=int, default=4)
    ap.add_argument("--learn-iters", type=int, default=50)
    ap.add_argument("--train-steps", type=int, default=1500)
    ap.add_argument("--batched-steps", type=int, default=3,
                    help="optimisation iterations of the 16-load-case batched workload (0: skip)")
    ap.add_argument("--profile-steps", type=int, default=2,
                    help="untimed iterations with per-kernel HIP-event timing (roofline)")
    ap.add_argument("--sweep-mode", type=int, default=3,
                    help="fine GS schedule: 0 colour passes, 1 colour-slab wavefront, 2 row pairs, "
                         "3 in-wave pairs (FP32 cycles)")
    ap.add_argument("--no-graph-replay", action="store_true",
                    help="launch the FCG iteration kernel by kernel (for launch-tracing tools)")
    ap.add_argument("--learn-only", action="store_true",
                    help="only generate the pairs and train the learned correction (network profiling)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only run a device-to-device copy of this size (HBM counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy_mib:
        import torch
        n = args.calibrate_copy_mib * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 5}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    from topoptspline3d_amd.backend import hip_api
    from topoptspline3d_amd import learned as LN
    net, learn_info = None, None
    if (not args.no_learned or args.learn_only) and args.workload == "c2":
        net, ds, learn_info = learned_phase(hip_api(), args.learn_domains, args.learn_iters, args.train_steps)
        ds.close()
        if args.learn_only:
            print(json.dumps({"learned": learn_info}), flush=True)
            return
    t0 = time.perf_counter()
    wl = make_workload(args.workload)
    log(f"[bench] workload {wl['name']} generated in {time.perf_counter() - t0:.1f}s {wl['meta'].get('fill', '')}")
    api = hip_api()

    def run(tag, precision, use_net, steps, keep=False, recycle=0):
        """warmup + steps iterations from the initial design; returns (loop, ms/iter, its, max res)"""
        loop = Loop(api, wl, precision, recycle)
        loop.ctx.set_sweep_mode(args.sweep_mode)
        loop.ctx.set_graph_replay(not args.no_graph_replay)
        if use_net:
            LN.attach(loop.ctx, net, level=2)
        for i in range(args.warmup):
            t = time.perf_counter(); loop.step()
            log(f"[bench] {tag} warmup {i}: {1e3 * (time.perf_counter() - t):.1f} ms, pcg its {loop.its[-1]}")
        sync()
        t = time.perf_counter()
        for i in range(steps):
            loop.step()
            log(f"[bench] {tag} step {i}: pcg its {loop.its[-1]} rel_res {loop.res[-1]:.2e} c {loop.c[-1]:.6e}")
        sync()
        ms = (time.perf_counter() - t) * 1e3 / max(steps, 1)
        n0 = args.warmup
        out = (loop, ms, loop.its[n0:], max(loop.res[n0:]) if steps else None)
        if not keep:
            if use_net:
                LN.attach(loop.ctx, None)
            loop.ctx.close()
        return out

    if args.probe_element_mfma:
        loop, *_ = run("plain", args.precision, False, 0, keep=True, recycle=args.recycle)
        ne = loop.ctx.info()["n_active_elements"][0]
        pr = loop.ctx.probe_element_mfma(reps=10)
        # useful work: 576 FMAs per element (the MFMA tiles execute 32/24 of it and more
        # where a row run has fewer than 16 elements)
        pr["useful_tflops"] = 2 * 576 * ne / (pr["mfma_products_ms"] * 1e-3) / 1e12
        pr["valu_useful_tflops"] = 2 * 576 * ne / (pr["valu_ms"] * 1e-3) / 1e12
        pr["active_elements"] = ne
        print(json.dumps({"element_mfma": pr}), flush=True)
        loop.ctx.close()
        return
    if args.surface_only:
        loop, *_ = run("plain", args.precision, False, 0, keep=True, recycle=args.recycle)
        print(json.dumps({"surface_constraints": surface_phase(loop.ctx, cells_list=args.surface_cells)}), flush=True)
        loop.ctx.close()
        return
    # headline: mixed-precision MG-PCG with the recycled start
    loop, ms, its, res_max = run("plain", args.precision, False, args.steps, keep=True, recycle=args.recycle)
    info = loop.ctx.info()
    nv = active_nodes(wl["solid"])
    log(f"[bench] levels {info['n_levels']} dims {info['level_dims']} coarse dofs {info['coarse_dofs']}")
    # per-kernel HIP-event timing in separate, untimed iterations (two events per launch
    # would otherwise add host work to the timed region)
    kernel_stats(api, loop.ctx, True)
    for i in range(args.profile_steps):
        loop.step()
        log(f"[bench] profiled step {i}: pcg its {loop.its[-1]}")
    sync()
    stats = kernel_stats(api, loop.ctx, False)
    line = {
        "metric": METRIC, "value": round(ms, 3), "unit": "ms/TO-iter", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPES[args.precision],
        "data": "synthetic",
        "config": {"workload": wl["desc"], "grid": list(wl["solid"].shape[::-1]),
                   "active_elements": info["n_active_elements"][0], "active_nodes": nv,
                   "mg_levels": info["n_levels"], "rtol": wl["rtol"], "precision": args.precision,
                   "solver": "FCG, " + SOLVERS[args.precision] + ", FP64 refinement (no learned correction)"
                             + (f", start projected onto the warm start and {args.recycle - 1} earlier solutions"
                                if args.recycle > 1 else ""),
                   "recycle": args.recycle,
                   "iterations_timed": f"{args.warmup}..{args.warmup + args.steps - 1} from the initial design",
                   "pcg_iterations": its, "max_rel_res": res_max},
        "roofline": roofline(stats, info, nv, args.profile_steps,
                             "ellipsoids_511x320x320" if wl["name"] == "c2" else None, args.precision,
                             args.sweep_mode),
    }
    ms_plain, its_plain = ms, its
    if args.recycle > 0 and not args.no_recycle_compare:
        # the same mixed-precision solver from the plain warm start (attribution)
        _, ms_plain, its_plain, res_plain = run("no-recycle", args.precision, False, args.steps)
        line["without_recycling"] = {"ms_per_iter": round(ms_plain, 3), "pcg_iterations": its_plain,
                                     "max_rel_res": res_plain, "speedup_of_recycling": round(ms_plain / ms, 3)}
    if not args.no_fp64_compare:
        # plain FP64 GPU multigrid (plain warm start) over the same iterations from the same
        # initial design
        _, ms64, its64, res64 = run("fp64", 0, False, args.steps)
        line["plain_fp64"] = {"ms_per_iter": round(ms64, 3), "pcg_iterations": its64, "max_rel_res": res64,
                              "speedup_of_value": round(ms64 / ms, 3),
                              "speedup_of_precision_alone": round(ms64 / ms_plain, 3)}
    if net is not None:
        # the learned correction runs in the FP64-residual FCG (precision 3 takes none)
        lloop, msl, itsl, resl = run("learned", min(args.precision, 2), True, args.steps, keep=True)
        tried, kept, stepk = LN.stats(lloop.ctx)
        LN.attach(lloop.ctx, None)
        lloop.ctx.close()
        line["learned"] = dict(learn_info, corrections_tried=tried, corrections_kept=kept,
                               mean_step_kept=round(stepk, 4), ms_per_iter=round(msl, 3), pcg_iterations=itsl,
                               max_rel_res=resl, speedup_over_plain=round(ms_plain / msl, 3),
                               iterations_saved=int(sum(its_plain) - sum(itsl)))
    if not args.no_fp64_compare:
        # same operator, solve from zero to the same tolerance: FP64, mixed, learned
        # (an OC step changes the densities themselves: rebuild the operator first)
        if not wl["spline"]:
            loop.ctx.update_operator()
        loop.ctx.set_start_recycling(0)  # from zero means from zero: no projected start
        cmp = {}
        modes = [("fp64", 0, False), ("mixed_fp32_stencils", 1, False), ("mixed_fp16_stencils", 2, False),
                 ("mixed_fp32_residual", 3, False)]
        if net is not None:
            modes.append(("learned", min(args.precision, 2), True))
        for tag, prec, use_net in modes:
            if use_net:
                LN.attach(loop.ctx, net, level=2)
            loop.ctx.set_displacement(np.zeros(3 * loop.ctx.nn))
            st = loop.ctx.solve(rtol=wl["rtol"], maxit=400, precision=prec, warm_start=True)
            if use_net:
                tried, kept, _ = LN.stats(loop.ctx)
                LN.attach(loop.ctx, None)
            cmp[tag] = {"solve_ms": round(st["seconds"] * 1e3, 2), "iterations": st["iterations"][0],
                        "rel_res": st["rel_res"][0]}
            if use_net:
                cmp[tag].update(corrections_tried=tried, corrections_kept=kept)
        line["solve_from_zero"] = cmp
    if not args.no_worst:
        if not wl["spline"]:
            loop.ctx.update_operator()
        line["worst_case_load"] = worst_phase(loop.ctx, wl, min(args.precision, 2))  # V-cycle precision
    coeffs_final = None
    if wl["spline"] and not args.no_surface:
        line["surface_constraints"] = surface_phase(loop.ctx, cells_list=args.surface_cells)
        coeffs_final = loop.ctx.coeffs()
    loop.ctx.close()
    if args.batched_steps > 0 and args.workload == "c2":
        t = time.perf_counter()
        line["batched_k16"] = batched_phase(api, args.batched_steps)
        log(f"[bench] batched workload took {time.perf_counter() - t:.1f}s")
    if not args.no_cpu_baseline:
        t = time.perf_counter()
        line["cpu_baseline"] = cpu_baseline(wl, its, coeffs_final)
        log(f"[bench] cpu baseline took {time.perf_counter() - t:.1f}s")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
