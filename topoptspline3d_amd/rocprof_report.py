"""Summaries of rocprofv3 CSV output for the roofline report.

* kernel time table from ``<name>_kernel_stats.csv`` (``rocprofv3 --kernel-trace --stats
  --output-format csv``), with a regex aggregate (calls, mean duration);
* HBM traffic per dispatch from two counter runs (``--pmc FETCH_SIZE`` and
  ``--pmc WRITE_SIZE``, KiB per dispatch), each scaled by what the same counter reads for
  a device copy of known size (``bench.py --calibrate-copy-mib N`` under the same
  counter), because the counters need not count every byte once.

A kernel class can consist of several kernels launched together (repeat --kernel, e.g.
the paired GS pass and the copy of its deferred nodes): their per-dispatch bytes and mean
durations are added.  --min-grid keeps only the counter dispatches with at least that
many work-items (a kernel launched on several multigrid levels: the largest level's
launches); with --trace (a `<name>_kernel_trace.csv`) the class's mean durations come from
the traced dispatches under the same filter, else from the per-name stats.

* issue rates from SQ/GRBM counter runs (``--compute``): VALU busy fraction, FP32 FMA
  share and rate, MFMA busy fraction per kernel (``compute_rates`` has the formulas).

Usage:
  python -m topoptspline3d_amd.rocprof_report --compute [--min-grid G] PATTERN... -- P1.csv P2.csv ...
  python -m topoptspline3d_amd.rocprof_report --stats DIR/c2_kernel_stats.csv \\
      --kernel "k_fine_gs<float" [--kernel ...] --fetch F.csv --write W.csv \\
      --cal-fetch CF.csv --cal-write CW.csv --copy-mib 1024 [--json out.json] [--meta k=v]
"""
import argparse
import csv
import json
import re
import sys


def _short(name):
    return re.sub(r"\(.*", "", name).replace("void ", "").replace("tos::", "")


def kernel_table(stats_csv, top=30):
    rows = list(csv.DictReader(open(stats_csv)))
    tot = sum(float(r["TotalDurationNs"]) for r in rows)
    out = ["| kernel | calls | total ms | avg us | % |", "|---|---:|---:|---:|---:|"]
    for r in rows[:top]:
        out.append(f"| `{_short(r['Name'])}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.2f} | "
                   f"{float(r['AverageNs']) / 1e3:.2f} | {float(r['Percentage']):.2f} |")
    out.append(f"\nall kernels: {len(rows)} names, {tot / 1e6:.1f} ms total")
    return "\n".join(out)


def kernel_mean(stats_csv, pattern):
    """(calls, mean duration in us) over the kernels whose name starts with `pattern`"""
    calls, ns = 0, 0.0
    for r in csv.DictReader(open(stats_csv)):
        if _short(r["Name"]).startswith(pattern):
            calls += int(r["Calls"])
            ns += float(r["TotalDurationNs"])
    return calls, (ns / calls / 1e3 if calls else None)


def _grid(r):
    """work-items of a dispatch row (counter CSVs: Grid_Size; trace CSVs: Grid_Size_X/Y/Z)"""
    if r.get("Grid_Size"):
        return int(r["Grid_Size"])
    return int(r.get("Grid_Size_X", 0) or 0) * int(r.get("Grid_Size_Y", 1) or 1) * int(r.get("Grid_Size_Z", 1) or 1)


def trace_mean(trace_csv, pattern, min_grid=0):
    """(dispatches, mean duration in us) of the traced dispatches of kernels matching
    `pattern` with at least `min_grid` work-items"""
    n, ns = 0, 0.0
    for r in csv.DictReader(open(trace_csv)):
        if _short(r["Kernel_Name"]).startswith(pattern) and _grid(r) >= min_grid:
            n += 1
            ns += float(r["End_Timestamp"]) - float(r["Start_Timestamp"])
    return n, (ns / n / 1e3 if n else None)


def counter_per_dispatch(counter_csv, pattern, min_grid=0):
    """(dispatches, mean counter value) over the dispatches of kernels matching `pattern`
    with at least `min_grid` work-items"""
    n, v = 0, 0.0
    for r in csv.DictReader(open(counter_csv)):
        if _short(r["Kernel_Name"]).startswith(pattern) and _grid(r) >= min_grid:
            n += 1
            v += float(r["Counter_Value"])
    return n, (v / n if n else None)


def calibration(counter_csv, copy_bytes):
    """bytes the copy moved / bytes the counter reported (counter in KiB)"""
    n, v = counter_per_dispatch(counter_csv, "__amd_rocclr_copyBuffer")
    if not n:
        raise SystemExit(f"no copy dispatch in {counter_csv}")
    return copy_bytes / (v * 1024.0)


def traffic(fetch_csv, write_csv, cal_fetch_csv, cal_write_csv, copy_mib, pattern, min_grid=0):
    copy_bytes = copy_mib * (1 << 20)
    kf = calibration(cal_fetch_csv, copy_bytes)
    kw = calibration(cal_write_csv, copy_bytes)
    nf, f = counter_per_dispatch(fetch_csv, pattern, min_grid)
    nw, w = counter_per_dispatch(write_csv, pattern, min_grid)
    return {"kernel": pattern, "dispatches": [nf, nw], "fetch_kib_raw": f, "write_kib_raw": w,
            "fetch_scale": kf, "write_scale": kw,
            "bytes_per_launch": f * 1024.0 * kf + w * 1024.0 * kw}


def counters(counter_csvs, pattern, min_grid=0):
    """{counter: mean value per dispatch} and the mean dispatch duration (us) over the
    dispatches of kernels matching `pattern` with at least `min_grid` work-items, from one
    or more counter runs (each run one pass of counters)"""
    vals, durs = {}, []
    for f in counter_csvs:
        for r in csv.DictReader(open(f)):
            if _short(r["Kernel_Name"]).startswith(pattern) and _grid(r) >= min_grid:
                vals.setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
                durs.append(float(r["End_Timestamp"]) - float(r["Start_Timestamp"]))
    mean = {k: sum(v) / len(v) for k, v in vals.items()}
    return mean, {k: len(v) for k, v in vals.items()}, (sum(durs) / len(durs) / 1e3 if durs else None)


def compute_rates(c, us, n_simd, n_xcd):
    """Derived issue rates from SQ/GRBM counters of one dispatch (means):
    gpu_cycles = GRBM_GUI_ACTIVE / n_xcd (the counter sums the XCDs);
    valu_busy = SQ_ACTIVE_INST_VALU * 4 / (n_simd * gpu_cycles) (quad-cycles per SIMD);
    fma_share = SQ_INSTS_VALU_FMA_F32 / SQ_INSTS_VALU; fp32_fma_tflops = 128 flops per
    wave-wide FMA instruction over the dispatch duration; mfma_busy =
    SQ_VALU_MFMA_BUSY_CYCLES / (n_simd * gpu_cycles)."""
    out = {}
    cyc = c.get("GRBM_GUI_ACTIVE")
    if cyc:
        cyc /= n_xcd
        out["gpu_cycles"] = cyc
        out["clock_ghz"] = cyc / (us * 1e3) if us else None
        if "SQ_ACTIVE_INST_VALU" in c:
            out["valu_busy"] = c["SQ_ACTIVE_INST_VALU"] * 4 / (n_simd * cyc)
        if "SQ_VALU_MFMA_BUSY_CYCLES" in c:
            out["mfma_busy"] = c["SQ_VALU_MFMA_BUSY_CYCLES"] / (n_simd * cyc)
    if "SQ_INSTS_VALU" in c and "SQ_INSTS_VALU_FMA_F32" in c:
        out["fma_share"] = c["SQ_INSTS_VALU_FMA_F32"] / c["SQ_INSTS_VALU"]
    if "SQ_INSTS_VALU" in c and "SQ_WAVES" in c:
        out["valu_insts_per_wave"] = c["SQ_INSTS_VALU"] / c["SQ_WAVES"]
    if "SQ_INSTS_VALU_FMA_F32" in c and us:
        out["fp32_fma_tflops"] = c["SQ_INSTS_VALU_FMA_F32"] * 128 / (us * 1e-6) / 1e12
    return out


def compute_table(counter_csvs, patterns, min_grid=0, n_simd=1024, n_xcd=8):
    """markdown table of compute_rates per kernel name (exact template instances)"""
    names = []
    for f in counter_csvs:
        for r in csv.DictReader(open(f)):
            k = _short(r["Kernel_Name"])
            if any(k.startswith(p) for p in patterns) and _grid(r) >= min_grid and k not in names:
                names.append(k)
    out = ["| kernel | dispatches | mean us | VALU busy | FP32 FMA share | VALU insts / wave | FP32 FMA TFLOP/s "
           "| MFMA busy | clock GHz |", "|---|---:|---:|---:|---:|---:|---:|---:|---:|"]
    for k in names:
        c, n, us = _exact(counter_csvs, k, min_grid)
        d = compute_rates(c, us, n_simd, n_xcd)
        f = lambda key, fmt: (fmt % d[key]) if d.get(key) is not None else "-"
        out.append(f"| `{k}` | {max(n.values())} | {us:.2f} | {f('valu_busy', '%.3f')} | {f('fma_share', '%.3f')} | "
                   f"{f('valu_insts_per_wave', '%.0f')} | {f('fp32_fma_tflops', '%.1f')} | {f('mfma_busy', '%.4f')} | "
                   f"{f('clock_ghz', '%.2f')} |")
    return "\n".join(out)


def _exact(counter_csvs, name, min_grid):
    vals, durs = {}, []
    for f in counter_csvs:
        for r in csv.DictReader(open(f)):
            if _short(r["Kernel_Name"]) == name and _grid(r) >= min_grid:
                vals.setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
                durs.append(float(r["End_Timestamp"]) - float(r["Start_Timestamp"]))
    return ({k: sum(v) / len(v) for k, v in vals.items()}, {k: len(v) for k, v in vals.items()},
            sum(durs) / len(durs) / 1e3 if durs else None)


def main(argv=None):
    if argv is None:
        argv = sys.argv[1:]
    if argv and argv[0] == "--compute":
        # python -m topoptspline3d_amd.rocprof_report --compute [--min-grid G] PATTERN... -- CSV...
        rest = argv[1:]
        min_grid = 0
        if rest[:1] == ["--min-grid"]:
            min_grid, rest = int(rest[1]), rest[2:]
        i = rest.index("--")
        print(compute_table(rest[i + 1:], rest[:i], min_grid))
        return 0
    ap = argparse.ArgumentParser()
    ap.add_argument("--stats", required=True)
    ap.add_argument("--kernel", required=True, action="append")
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--cal-fetch")
    ap.add_argument("--cal-write")
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--min-grid", type=int, default=0, help="dispatches with at least this many work-items")
    ap.add_argument("--trace", help="kernel trace CSV for the class's mean durations")
    ap.add_argument("--top", type=int, default=30)
    ap.add_argument("--json")
    ap.add_argument("--meta", action="append", default=[], help="key=value recorded in the JSON")
    a = ap.parse_args(argv)
    print(kernel_table(a.stats, a.top))
    names = " + ".join(f"`{k}*`" for k in a.kernel)
    calls, us = 0, 0.0
    for k in a.kernel:
        ck, uk = trace_mean(a.trace, k, a.min_grid) if a.trace else kernel_mean(a.stats, k)
        print(f"\n`{k}*`: {ck} launches, mean {uk:.2f} us")
        calls = ck if not calls else min(calls, ck)
        us += uk
    out = {"kernel": a.kernel if len(a.kernel) > 1 else a.kernel[0], "launches": calls, "mean_us": us}
    if a.min_grid:
        out["min_grid"] = a.min_grid
    for kv in a.meta:
        k, v = kv.split("=", 1)
        out[k] = int(v) if v.lstrip("-").isdigit() else v
    if a.fetch:
        parts = [traffic(a.fetch, a.write, a.cal_fetch, a.cal_write, a.copy_mib, k, a.min_grid) for k in a.kernel]
        t = dict(parts[0]) if len(parts) == 1 else {"kernel": a.kernel, "parts": parts,
                                                    "bytes_per_launch": sum(p["bytes_per_launch"] for p in parts)}
        out["traffic"] = t
        for p in parts:
            print(f"\nHBM traffic per launch of `{p['kernel']}*`: FETCH_SIZE {p['fetch_kib_raw']:.0f} KiB x "
                  f"{p['fetch_scale']:.3f} + WRITE_SIZE {p['write_kib_raw']:.0f} KiB x {p['write_scale']:.3f} = "
                  f"{p['bytes_per_launch'] / 1e6:.1f} MB ({p['dispatches'][0]} / {p['dispatches'][1]} dispatches)")
        if len(parts) > 1:
            print(f"\nclass {names}: {t['bytes_per_launch'] / 1e6:.1f} MB per launch, mean {us:.2f} us")
    if a.json:
        json.dump(out, open(a.json, "w"), indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
