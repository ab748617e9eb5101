"""Summaries of rocprofv3 CSV output for the sweep kernels.

  python -m multistageqw_amd.profiling stats <dir>              kernel-trace --stats table
  python -m multistageqw_amd.profiling pmc <dir> <counter> <kernel substring> [min workgroups]
        per-dispatch counter values summed over counter dimensions
  python -m multistageqw_amd.profiling trace <dir> <kernel substring> [min workgroups]
        dispatch count and average duration from a --kernel-trace run

A minimum workgroup count keeps the term sweeps of a batch and drops the
small launches of the same kernel (the shared first-stage basis recursion
runs one job per instance).

FETCH_SIZE / WRITE_SIZE readings are in KiB per dispatch; on MI355X the
calibration copy (hbm_copy) reads FETCH_SIZE = half the bytes it moves and
WRITE_SIZE = all of them, so HBM bytes = 2 * FETCH_SIZE + WRITE_SIZE (x 1024).
"""
from __future__ import annotations

import collections
import csv
import glob
import os
import sys


def _one(d, suffix):
    files = sorted(glob.glob(os.path.join(d, f"*{suffix}")))
    if not files:
        raise FileNotFoundError(f"no *{suffix} under {d}")
    return files[0]


def stats(d):
    rows = list(csv.DictReader(open(_one(d, "kernel_stats.csv"))))
    return [(r["Name"], int(r["Calls"]), float(r["TotalDurationNs"]), float(r["AverageNs"]), float(r["Percentage"]))
            for r in rows]


def pmc(d, counter, kernel, min_wgs=0):
    rows = list(csv.DictReader(open(_one(d, "counter_collection.csv"))))
    per = collections.OrderedDict()
    for r in rows:
        if r["Counter_Name"] != counter or kernel not in r["Kernel_Name"]:
            continue
        did = int(r["Dispatch_Id"])
        g = int(r["Grid_Size"]) // max(1, int(r["Workgroup_Size"]))
        if g < min_wgs:
            continue
        v, _ = per.get(did, (0.0, g))
        per[did] = (v + float(r["Counter_Value"]), g)
    return [(did, v, g) for did, (v, g) in per.items()]


def trace(d, kernel, min_wgs=0):
    """(count, average ns, total ns) of the dispatches of a kernel with at least min_wgs workgroups."""
    durs = []
    for r in csv.DictReader(open(_one(d, "kernel_trace.csv"))):
        if kernel not in r["Kernel_Name"]:
            continue
        if int(r["Grid_Size_X"]) // max(1, int(r["Workgroup_Size_X"])) < min_wgs:
            continue
        durs.append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    return len(durs), (sum(durs) / len(durs) if durs else 0.0), float(sum(durs))


def main(argv):
    if len(argv) >= 2 and argv[0] == "stats":
        for name, calls, tot, avg, pct in stats(argv[1]):
            print(f"{pct:6.2f}%  calls {calls:6d}  avg {avg / 1e3:10.1f} us  total {tot / 1e6:9.2f} ms  {name[:90]}")
        return 0
    if len(argv) >= 4 and argv[0] == "pmc":
        vals = pmc(argv[1], argv[2], argv[3], int(argv[4]) if len(argv) > 4 else 0)
        tot = sum(v for _, v, _ in vals)
        wgs = sum(g for _, _, g in vals)
        print(f"{argv[2]} over {len(vals)} dispatches of '{argv[3]}': total {tot:.1f}, per dispatch "
              f"{tot / max(1, len(vals)):.1f}, per workgroup {tot / max(1, wgs):.3f}, workgroups {wgs}")
        return 0
    if len(argv) >= 3 and argv[0] == "trace":
        cnt, avg, tot = trace(argv[1], argv[2], int(argv[3]) if len(argv) > 3 else 0)
        print(f"{cnt} dispatches of '{argv[2]}': average {avg / 1e3:.1f} us, total {tot / 1e6:.2f} ms")
        return 0
    print(__doc__)
    return 2


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
