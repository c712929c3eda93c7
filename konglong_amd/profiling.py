"""Summarise rocprofv3 CSV output (kernel stats and FETCH_SIZE / WRITE_SIZE passes).

Usage:
  python -m konglong_amd.profiling --stats DIR --fetch DIR --write DIR --group PREFIX[,PREFIX...] \
      --title TEXT --command TEXT > profiles/<kernel>_<workload>.txt

FETCH_SIZE and WRITE_SIZE (KiB per dispatch) are scaled by the ratio of the
known bytes of a plain device-to-device copy in the same run (the largest
__amd_rocclr_copyBuffer dispatch, `copy_bytes` read and written) to what the
counter reports for it.  The per-step totals of the kernels whose names start
with one of the group prefixes (each weighted by its launches per step) are written as `fetch_bytes_per_step` and
`write_bytes_per_step` lines, which bench.py reads as the roofline traffic.  With a kernel trace in the stats
directory, `step_kernel_us` is the time per step during which at least one group kernel runs (kernels on
concurrent streams count once); the plain sum is `step_kernel_sum_us`.
"""
import argparse
import csv
import glob
import os
import re
from collections import defaultdict


def _short(name):
    m = re.match(r"(?:void )?(?:klrtm::)?([A-Za-z0-9_]+)(<[^>]*>)?", name)
    return (m.group(1) + (m.group(2) or "")) if m else name[:60]


def read_stats(d):
    f = glob.glob(os.path.join(d, "*kernel_stats.csv"))[0]
    out = {}
    for r in csv.DictReader(open(f)):
        out[_short(r["Name"])] = (int(r["Calls"]), float(r["AverageNs"]) / 1e3, float(r["Percentage"]))
    return out


def read_busy(d, prefixes):
    """Time (ns) during which at least one kernel of the group runs, from the kernel trace: the union of the
    group's dispatch intervals, so that concurrent kernels count once (None without a trace)."""
    fs = glob.glob(os.path.join(d, "*kernel_trace.csv"))
    if not fs:
        return None
    iv = []
    for r in csv.DictReader(open(fs[0])):
        if any(_short(r["Kernel_Name"]).startswith(p) for p in prefixes):
            iv.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"])))
    iv.sort()
    busy, cur0, cur1 = 0, None, None
    for a, b in iv:
        if cur1 is None or a > cur1:
            if cur1 is not None:
                busy += cur1 - cur0
            cur0, cur1 = a, b
        else:
            cur1 = max(cur1, b)
    if cur1 is not None:
        busy += cur1 - cur0
    return busy


def read_counter(d, name):
    f = glob.glob(os.path.join(d, "*counter_collection.csv"))[0]
    vals = defaultdict(list)
    grid = defaultdict(list)
    for r in csv.DictReader(open(f)):
        if r["Counter_Name"] != name:
            continue
        k = _short(r["Kernel_Name"])
        vals[k].append(float(r["Counter_Value"]))
        grid[k].append(int(r["Grid_Size"]))
    return vals, grid


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--stats", required=True)
    ap.add_argument("--fetch", required=True)
    ap.add_argument("--write", required=True)
    ap.add_argument("--group", required=True, help="comma separated kernel-name prefixes of one step")
    ap.add_argument("--copy-bytes", type=float, required=True, help="bytes read (= written) by the largest copy")
    ap.add_argument("--title", default="")
    ap.add_argument("--command", action="append", default=[])
    a = ap.parse_args()
    stats = read_stats(a.stats)
    fv, fg = read_counter(a.fetch, "FETCH_SIZE")
    wv, wg = read_counter(a.write, "WRITE_SIZE")

    def largest_copy(vals, grid):
        k = [n for n in vals if n.startswith("__amd_rocclr_copyBuffer")]
        if not k:
            return None
        pairs = sorted(zip(grid[k[0]], vals[k[0]]))
        return pairs[-1][1] * 1024.0
    cf, cw = largest_copy(fv, fg), largest_copy(wv, wg)
    kf = a.copy_bytes / cf if cf else 1.0
    kw = a.copy_bytes / cw if cw else 1.0
    prefixes = a.group.split(",")
    print(f"# {a.title}")
    for c in a.command:
        print(f"# command: {c}")
    print(f"# counter scale: FETCH_SIZE x{kf:.3f}, WRITE_SIZE x{kw:.3f} "
          f"(largest device copy of {a.copy_bytes / 1e6:.1f} MB read and written)")
    print(f"{'kernel':60s} {'calls':>6s} {'avg_us':>10s} {'pct':>6s} {'fetch_MB':>10s} {'write_MB':>10s}")
    # a kernel launched twice per step (the same template for two passes) has twice the calls of the
    # others in the group; the step count is the smallest call count in the group
    in_group = [k for k in stats if any(k.startswith(p) for p in prefixes)]
    steps = min(stats[k][0] for k in in_group) if in_group else 1
    tf = tw = tus = 0.0
    for k, (calls, us, pct) in sorted(stats.items(), key=lambda kv: -kv[1][2]):
        f = sum(fv[k]) / len(fv[k]) * 1024 * kf if fv.get(k) else float("nan")
        w = sum(wv[k]) / len(wv[k]) * 1024 * kw if wv.get(k) else float("nan")
        print(f"{k[:60]:60s} {calls:6d} {us:10.1f} {pct:6.2f} {f / 1e6:10.1f} {w / 1e6:10.1f}")
        if k in in_group:
            per = calls / steps
            tf += f * per
            tw += w * per
            tus += us * per
    busy = read_busy(a.stats, prefixes)
    if busy is not None:   # concurrent kernels of one step counted once
        print(f"step_kernel_sum_us: {tus:.1f}")
        tus = busy / 1e3 / steps
    print(f"step_kernel_us: {tus:.1f}")
    print(f"fetch_bytes_per_step: {tf:.0f}")
    print(f"write_bytes_per_step: {tw:.0f}")
    print(f"hbm_GBps_measured: {(tf + tw) / (tus * 1e-6) / 1e9:.1f}")


if __name__ == "__main__":
    main()
