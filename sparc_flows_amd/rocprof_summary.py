"""Summaries of rocprofv3 CSV output (``-f csv``), small enough to keep as text.

    python3 -m sparc_flows_amd.rocprof_summary stats DIR [--prune]
    python3 -m sparc_flows_amd.rocprof_summary pmc DIR [--prune]

``stats`` prints the per-kernel rows of ``*kernel_stats.csv`` (calls, total,
average, min and max duration) and, when a ``*kernel_trace.csv`` is present,
the number of dispatches, the summed kernel time, the time span they cover and
the idle gaps between back-to-back dispatches.
``pmc`` averages every counter of ``*counter_collection.csv`` per kernel over
its dispatches (the value is per dispatch, e.g. KiB for FETCH_SIZE and
WRITE_SIZE).  ``--prune`` deletes the per-dispatch CSV files and database files
afterwards, which are large for long runs.
"""
from __future__ import annotations

import argparse
import collections
import csv
import glob
import os
import sys


def _find(d: str, pat: str):
    f = sorted(glob.glob(os.path.join(d, "**", pat), recursive=True))
    return f[0] if f else None


def stats(d: str, out=sys.stdout, top: int = 16):
    f = _find(d, "*kernel_stats.csv")
    if f is None:
        raise SystemExit(f"no kernel_stats.csv under {d}")
    with open(f) as fh:
        rows = list(csv.DictReader(fh))
    for r in rows[:top]:
        n = r.get("Name", "")[:72]
        print(f"{n:72s} calls {r.get('Calls')} total_ms {float(r.get('TotalDurationNs', 0)) / 1e6:.3f} "
              f"avg_us {float(r.get('AverageNs', 0)) / 1e3:.2f} min_us {float(r.get('MinNs', 0)) / 1e3:.2f} "
              f"max_us {float(r.get('MaxNs', 0)) / 1e3:.2f} pct {r.get('Percentage')}", file=out)
    t = _find(d, "*kernel_trace.csv")
    if t is not None:
        with open(t) as fh:
            st = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"])) for r in csv.DictReader(fh))
        if st:
            busy = sum(e - s for s, e in st)
            print(f"dispatches {len(st)} busy_ms {busy / 1e6:.1f} span_ms {(st[-1][1] - st[0][0]) / 1e6:.1f}",
                  file=out)
            # idle time between consecutive dispatches that follow each other
            # within 100 us (back-to-back launches, e.g. inside a graph)
            gaps = sorted(b[0] - a[1] for a, b in zip(st, st[1:]) if 0 <= b[0] - a[1] < 100_000)
            if gaps:
                print(f"back-to-back gaps {len(gaps)} median_us {gaps[len(gaps) // 2] / 1e3:.2f} "
                      f"mean_us {sum(gaps) / len(gaps) / 1e3:.2f}", file=out)
    return t


def pmc(d: str, out=sys.stdout):
    f = _find(d, "*counter_collection.csv")
    if f is None:
        raise SystemExit(f"no counter_collection.csv under {d}")
    agg = collections.defaultdict(list)
    with open(f) as fh:
        for r in csv.DictReader(fh):
            agg[(r["Kernel_Name"][:72], r["Counter_Name"])].append(float(r["Counter_Value"]))
    for (k, c), v in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        print(f"{k:72s} {c} dispatches {len(v)} mean {sum(v) / len(v):.1f} min {min(v):.1f} max {max(v):.1f}",
              file=out)
    return f


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("mode", choices=("stats", "pmc"))
    ap.add_argument("dir")
    ap.add_argument("--prune", action="store_true",
                    help="delete the per-dispatch CSV and database files afterwards")
    a = ap.parse_args(argv)
    big = stats(a.dir) if a.mode == "stats" else pmc(a.dir)
    if a.prune:
        for f in ([big] if big else []) + glob.glob(os.path.join(a.dir, "**", "*.db"), recursive=True):
            os.remove(f)
    return 0


if __name__ == "__main__":
    sys.exit(main())
