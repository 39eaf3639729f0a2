"""Summarise rocprofv3 runs of a sweep workload into a short text report.

    python -m girih_amd.profsum --trace T.db --fetch F.db --write W.db \
        --cfetch CF.db --cwrite CW.db --kernel sweep_kernel --lups N --algo-bytes B \
        --cmd "..." --title "..." --out profiles/<name>.md [--traffic-json profiles/traffic.json --key K]

Inputs are the rocpd SQLite databases rocprofv3 writes (`-d DIR -o NAME` gives
DIR/NAME_results.db): one --kernel-trace --stats run, one --pmc FETCH_SIZE run,
one --pmc WRITE_SIZE run of the workload, and the same two counters over
`python -m girih_amd.calib` (a streaming kernel of known size).  The counters
are turned into bytes with the factor the calibration kernel gives (bytes it
moved / counter reading), and reported per launch and per lattice update.
"""
import argparse
import json
import os
import re
import sqlite3
import statistics


def _short(name):
    return re.sub(r"\s+", " ", name)


def trace_rows(db):
    c = sqlite3.connect(db)
    return list(c.execute("select name, total_calls, total_duration, average, percentage from top_kernels "
                          "order by total_duration desc"))


def durations(db, kernel):
    c = sqlite3.connect(db)
    return [r[0] for r in c.execute("select duration from kernels where name like ? order by start",
                                    ("%" + kernel + "%",))]


def pmc(db, counter, kernel):
    """Counter value per dispatch (summed over the instances rocprofv3 reports)."""
    c = sqlite3.connect(db)
    per = {}
    for disp, name, val in c.execute("select dispatch_id, name, counter_value from pmc_events "
                                     "where counter_name = ?", (counter,)):
        if kernel in name:
            per[disp] = per.get(disp, 0.0) + float(val)
    return [per[k] for k in sorted(per)]


def calib_factor(db, counter, nbytes):
    vals = pmc(db, counter, "elementwise")
    if not vals:
        raise SystemExit("calibration kernel not found in %s" % db)
    return nbytes / (statistics.median(vals) * 1024.0), statistics.median(vals)


def main():
    ap = argparse.ArgumentParser()
    for k in ("trace", "fetch", "write", "cfetch", "cwrite", "kernel", "cmd", "title", "out"):
        ap.add_argument("--" + k, required=True)
    ap.add_argument("--lups", type=float, required=True, help="lattice updates per launch")
    ap.add_argument("--algo-bytes", type=float, required=True, help="algorithmic HBM bytes per launch")
    ap.add_argument("--calib-mib", type=int, default=1024)
    ap.add_argument("--traffic-json")
    ap.add_argument("--key")
    ap.add_argument("--skip", type=int, default=0, help="leading dispatches of the kernel to ignore")
    ap.add_argument("--steps-per-launch", type=int, default=1)
    ap.add_argument("--launch-config", default=None, help="the launch configuration profiled (bench.py --config)")
    a = ap.parse_args()

    rows = trace_rows(a.trace)
    d = durations(a.trace, a.kernel)[a.skip:]
    if not d:
        raise SystemExit("kernel %s not in trace" % a.kernel)
    avg_ns = statistics.mean(d)
    nb = a.calib_mib * 1024 * 1024
    sf, cf_raw = calib_factor(a.cfetch, "FETCH_SIZE", nb)
    sw, cw_raw = calib_factor(a.cwrite, "WRITE_SIZE", nb)
    f = pmc(a.fetch, "FETCH_SIZE", a.kernel)[a.skip:]
    w = pmc(a.write, "WRITE_SIZE", a.kernel)[a.skip:]
    fb = statistics.mean(f) * 1024.0 * sf
    wb = statistics.mean(w) * 1024.0 * sw
    algo_gbs = a.algo_bytes / (avg_ns * 1e-9) / 1e9
    hbm_gbs = (fb + wb) / (avg_ns * 1e-9) / 1e9
    lines = [
        "# %s" % a.title,
        "",
        "Commands (run from the repository root; rocprofv3 writes rocpd SQLite databases):",
        "",
        "```",
        a.cmd.strip(),
        "```",
        "",
        "## Kernel trace (--kernel-trace --stats), top kernels by total time",
        "",
        "| kernel | calls | avg ms | % time |",
        "|---|---|---|---|",
    ]
    for name, calls, tot, avg, pct in rows[:8]:
        lines.append("| `%s` | %d | %.3f | %.1f |" % (_short(name)[:110], calls, avg / 1e3, pct))
    lines += [
        "",
        "## `%s` (%d dispatches timed%s)" % (a.kernel, len(d), ", first %d skipped" % a.skip if a.skip else ""),
        "",
        "- average duration: %.2f us (min %.2f, max %.2f)" % (avg_ns / 1e3, min(d) / 1e3, max(d) / 1e3),
        "- lattice updates per launch: %.0f -> %.1f GLUP/s" % (a.lups, a.lups / (avg_ns * 1e-9) / 1e9),
        "- algorithmic bytes per launch: %.4g B (%.2f B/LUP) -> %.0f GB/s = %.3f of 8000 GB/s" % (
            a.algo_bytes, a.algo_bytes / a.lups, algo_gbs, algo_gbs / 8000.0),
        "",
        "## HBM traffic (--pmc FETCH_SIZE, --pmc WRITE_SIZE; %d dispatches each)" % len(f),
        "",
        "- calibration (`python -m girih_amd.calib`, %d MiB read + %d MiB written per dispatch): "
        "FETCH_SIZE %.1f KiB -> factor %.4f, WRITE_SIZE %.1f KiB -> factor %.4f" % (
            a.calib_mib, a.calib_mib, cf_raw, sf, cw_raw, sw),
        "- fetched: %.4g B per launch (%.2f B/LUP)" % (fb, fb / a.lups),
        "- written: %.4g B per launch (%.2f B/LUP)" % (wb, wb / a.lups),
        "- total: %.4g B per launch (%.2f B/LUP, %.2fx the algorithmic bytes) -> %.0f GB/s measured HBM" % (
            fb + wb, (fb + wb) / a.lups, (fb + wb) / a.algo_bytes, hbm_gbs),
        "",
    ]
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    with open(a.out, "w") as fh:
        fh.write("\n".join(lines))
    print("\n".join(lines))
    if a.traffic_json and a.key:
        data = {}
        if os.path.exists(a.traffic_json):
            with open(a.traffic_json) as fh:
                data = json.load(fh)
        data[a.key] = {"bytes_per_launch": fb + wb, "fetch_bytes": fb, "write_bytes": wb,
                       "bytes_per_lup": (fb + wb) / a.lups, "avg_launch_us": avg_ns / 1e3,
                       "steps_per_launch": a.steps_per_launch, "launch_config": a.launch_config,
                       "summary": os.path.relpath(a.out, os.path.dirname(os.path.abspath(a.traffic_json)))}
        with open(a.traffic_json, "w") as fh:
            json.dump(data, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
