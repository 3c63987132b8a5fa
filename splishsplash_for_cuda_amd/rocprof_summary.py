"""Summarise the rocprofv3 output of rocprof_passes.sh.

    python splishsplash_for_cuda_amd/rocprof_summary.py OUTDIR [--config c2] [--write-profiles]

Prints (and with --write-profiles writes under profiles/):
  * the kernel-trace statistics (calls, average and total duration per kernel);
  * HBM bytes per dispatch of each solver sweep from FETCH_SIZE and WRITE_SIZE
    (KiB per dispatch), each scaled by the factor measured on the 256 MiB copy
    kernel (expected / measured);
  * VALUBusy per sweep when that pass ran.
profiles/traffic_<sweep>_<config>.json carries the bytes per launch that
bench.py reports as roofline.traffic.
"""
from __future__ import annotations

import argparse
import csv
import json
import os
from collections import defaultdict

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# bench.py sweep name -> device symbol prefix (the last template argument of
# the sweeps selects the rigid-body variant; both count)
SWEEPS = {
    "div_ws_init": "void dfsph::k_div_ws_init<",
    "div_init": "void dfsph::k_div_source<true,",
    "div_jacobi": "void dfsph::k_jacobi<true,",
    "div_loop_end": "void dfsph::k_div_source<false,",
    "xsph": "dfsph::k_xsph(",
    "pressure_ws": "void dfsph::k_pressure_ws<",
    "dens_init": "void dfsph::k_density_source<true,",
    "dens_jacobi": "void dfsph::k_jacobi<false,",
    "dens_loop_end": "void dfsph::k_density_source<false,",
    "neighbors": "void dfsph::k_neighbors<",
}
COPY_BYTES = 256 * 1024 * 1024


def sweep_of(name):
    for k, pre in SWEEPS.items():
        if name.startswith(pre):
            return k
    return None


def short(name):
    return name.split("(")[0].replace("void ", "")


def read_counter(path, counter):
    """{kernel name: [values]} of one counter CSV."""
    out = defaultdict(list)
    if not os.path.exists(path):
        return out
    with open(path) as f:
        for row in csv.DictReader(f):
            if row["Counter_Name"] == counter:
                out[row["Kernel_Name"]].append(float(row["Counter_Value"]))
    return out


def min_working(vals, frac=0.5):
    """Values of working dispatches: velocity sweeps that exit at once (loop
    already converged) read almost nothing and are left out."""
    if not vals:
        return vals
    top = max(vals)
    return [v for v in vals if v >= frac * top]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("outdir")
    ap.add_argument("--config", default="c2")
    ap.add_argument("--write-profiles", action="store_true")
    ap.add_argument("--command", default="bash splishsplash_for_cuda_amd/rocprof_passes.sh OUTDIR --steps 5 "
                                         "--warmup 5 --no-cpu-baseline")
    ap.add_argument("--window", default=None,
                    help="W,K,P: the bench run's warm-up, timed and profiled steps; the kernel trace is then "
                         "also averaged over the plain leg's profiled steps (steps W+K+1 .. W+K+P), the "
                         "launches bench.py times for its roofline")
    args = ap.parse_args()
    d = args.outdir
    lines = []
    lines.append("# rocprofv3 summary: DFSPH step, config %s" % args.config.upper())
    lines.append("")
    lines.append("Command: `%s`" % args.command)
    lines.append("")

    stats_path = os.path.join(d, "trace", "run_kernel_stats.csv")
    if os.path.exists(stats_path):
        lines.append("## Kernel trace statistics (rocprofv3 --kernel-trace --stats)")
        lines.append("")
        lines.append("| kernel | calls | avg us | min us | total ms | % |")
        lines.append("|---|---|---|---|---|---|")
        with open(stats_path) as f:
            for row in csv.DictReader(f):
                lines.append("| %s | %s | %.1f | %.1f | %.2f | %.1f |" % (
                    short(row["Name"]), row["Calls"], float(row["AverageNs"]) / 1e3, float(row["MinNs"]) / 1e3,
                    float(row["TotalDurationNs"]) / 1e6, float(row["Percentage"])))
        lines.append("")
        lines.append("The velocity sweeps (`k_jacobi`) include at most one launch per solve that exits at once "
                     "(loop already converged); their minimum shows it.")
        lines.append("")

    trace_path = os.path.join(d, "trace", "run_kernel_trace.csv")
    if args.window and os.path.exists(trace_path):
        w, k, pst = (int(x) for x in args.window.split(","))
        with open(trace_path) as f:
            rows = sorted(((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"])
                           for r in csv.DictReader(f)), key=lambda r: r[0])
        ends = [e for (_, e, nm) in rows if nm.startswith("dfsph::k_update_pos(")]
        if len(ends) >= w + k + pst:
            t0, t1 = ends[w + k - 1], ends[w + k + pst - 1]
            per = defaultdict(list)
            for (a0, a1, nm) in rows:
                if t0 < a0 <= t1:
                    per[nm].append((a1 - a0) / 1e3)
            lines.append("## Profiled window (steps %d-%d of the first leg, the launches bench.py times)"
                         % (w + k + 1, w + k + pst))
            lines.append("")
            lines.append("| kernel | calls | avg us | min us |")
            lines.append("|---|---|---|---|")
            for nm, v in sorted(per.items(), key=lambda kv: -sum(kv[1]))[:16]:
                lines.append("| %s | %d | %.1f | %.1f |" % (short(nm), len(v), sum(v) / len(v), min(v)))
            lines.append("")

    # calibration factors
    factors = {}
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        vals = read_counter(os.path.join(d, "copy_" + c, "copy_counter_collection.csv"), c)
        copy_vals = [v for k, vs in vals.items() if "copy_f4" in k for v in vs]
        if copy_vals:
            measured = sum(copy_vals) / len(copy_vals) * 1024.0
            factors[c] = COPY_BYTES / measured
    if factors:
        lines.append("## Counter calibration (256 MiB copy kernel)")
        lines.append("")
        for c, f in factors.items():
            lines.append("- %s: expected %d B, factor expected/measured = %.4f" % (c, COPY_BYTES, f))
        lines.append("")

    fetch = read_counter(os.path.join(d, "bench_FETCH_SIZE", "bench_counter_collection.csv"), "FETCH_SIZE")
    write = read_counter(os.path.join(d, "bench_WRITE_SIZE", "bench_counter_collection.csv"), "WRITE_SIZE")
    valu = read_counter(os.path.join(d, "bench_VALUBusy", "bench_counter_collection.csv"), "VALUBusy")
    traffic = {}
    if fetch:
        lines.append("## HBM traffic per dispatch (calibrated)")
        lines.append("")
        lines.append("| sweep | dispatches | FETCH MB | WRITE MB | total MB | VALUBusy % |")
        lines.append("|---|---|---|---|---|---|")
        for name in sorted(fetch):
            sw = sweep_of(name)
            if not sw:
                continue
            fv = min_working(fetch[name])
            wv = min_working(write.get(name, []))
            fb = sum(fv) / len(fv) * 1024.0 * factors.get("FETCH_SIZE", 1.0)
            wb = (sum(wv) / len(wv) * 1024.0 * factors.get("WRITE_SIZE", 1.0)) if wv else 0.0
            vb = valu.get(name, [])
            vbs = "%.1f" % (sum(vb) / len(vb)) if vb else "-"
            traffic[sw] = {"bytes_per_launch": fb + wb, "fetch_bytes": fb, "write_bytes": wb, "dispatches": len(fv),
                           "kernel": short(name)}
            lines.append("| %s | %d | %.1f | %.1f | %.1f | %s |" % (sw, len(fv), fb / 1e6, wb / 1e6, (fb + wb) / 1e6, vbs))
        lines.append("")
    # further counter passes (EXTRA_COUNTERS): per sweep averages of every counter
    extra = sorted(e for e in os.listdir(d) if e.startswith("bench_") and e not in
                   ("bench_FETCH_SIZE", "bench_WRITE_SIZE", "bench_VALUBusy")) if os.path.isdir(d) else []
    for e in extra:
        path = os.path.join(d, e, "bench_counter_collection.csv")
        if not os.path.exists(path):
            continue
        per = defaultdict(lambda: defaultdict(list))
        with open(path) as f:
            for row in csv.DictReader(f):
                sw = sweep_of(row["Kernel_Name"])
                if sw:
                    per[sw][row["Counter_Name"]].append(float(row["Counter_Value"]))
        names = sorted({c for v in per.values() for c in v})
        if not names:
            continue
        lines.append("## Counters: %s (average per dispatch)" % e[len("bench_"):].replace("+", ", "))
        lines.append("")
        lines.append("| sweep | " + " | ".join(names) + " |")
        lines.append("|---|" + "---|" * len(names))
        for sw in SWEEPS:
            if sw in per:
                vals = [per[sw].get(c, []) for c in names]
                lines.append("| %s | " % sw + " | ".join("%.4g" % (sum(v) / len(v)) if v else "-" for v in vals) + " |")
        lines.append("")
    text = "\n".join(lines)
    print(text)
    if args.write_profiles:
        pdir = os.path.join(ROOT, "profiles")
        os.makedirs(pdir, exist_ok=True)
        with open(os.path.join(pdir, "dfsph_step_%s.md" % args.config), "w") as f:
            f.write(text + "\n")
        for sw, t in traffic.items():
            t = dict(t)
            t["command"] = args.command
            t["calibration"] = factors
            with open(os.path.join(pdir, "traffic_%s_%s.json" % (sw, args.config)), "w") as f:
                json.dump(t, f, indent=1)


if __name__ == "__main__":
    main()
