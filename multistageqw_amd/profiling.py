"""Summaries of rocprofv3 CSV output for the sweep kernels.

  python -m multistageqw_amd.profiling stats <dir>              kernel-trace --stats table
  python -m multistageqw_amd.profiling pmc <dir> <counter> <kernel substring>
        per-dispatch counter values summed over counter dimensions

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


def pmc(d, counter, kernel):
    rows = list(csv.DictReader(open(_one(d, "counter_collection.csv"))))
    per = collections.OrderedDict()
    for r in rows:
        if r["Counter_Name"] != counter or kernel not in r["Kernel_Name"]:
            continue
        did = int(r["Dispatch_Id"])
        g = int(r["Grid_Size"]) // max(1, int(r["Workgroup_Size"]))
        v, _ = per.get(did, (0.0, g))
        per[did] = (v + float(r["Counter_Value"]), g)
    return [(did, v, g) for did, (v, g) in per.items()]


def main(argv):
    if len(argv) >= 2 and argv[0] == "stats":
        for name, calls, tot, avg, pct in stats(argv[1]):
            print(f"{pct:6.2f}%  calls {calls:6d}  avg {avg / 1e3:10.1f} us  total {tot / 1e6:9.2f} ms  {name[:90]}")
        return 0
    if len(argv) >= 4 and argv[0] == "pmc":
        vals = pmc(argv[1], argv[2], argv[3])
        tot = sum(v for _, v, _ in vals)
        wgs = sum(g for _, _, g in vals)
        print(f"{argv[2]} over {len(vals)} dispatches of '{argv[3]}': total {tot:.1f}, per dispatch "
              f"{tot / max(1, len(vals)):.1f}, per workgroup {tot / max(1, wgs):.3f}")
        return 0
    print(__doc__)
    return 2


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
