"""Summaries of rocprofv3 CSV output (``-f csv``), small enough to keep as text.

    python3 -m sparc_flows_amd.rocprof_summary stats DIR [--prune] [--by-grid]
    python3 -m sparc_flows_amd.rocprof_summary pmc DIR [--prune] [--by-grid]

``stats`` prints the per-kernel rows of ``*kernel_stats.csv`` (calls, total,
average, min and max duration) and, when a ``*kernel_trace.csv`` is present,
the number of dispatches, the summed kernel time, the time span they cover and
the idle gaps between back-to-back dispatches.
``pmc`` averages every counter of ``*counter_collection.csv`` per kernel over
its dispatches (the value is per dispatch, e.g. KiB for FETCH_SIZE and
WRITE_SIZE).  ``--prune`` deletes the per-dispatch CSV files and database files
afterwards, which are large for long runs.  ``--by-grid`` keeps dispatches of
one kernel with different grid sizes apart (e.g. a one-source and a
16-source batch of the same instantiation); ``stats`` then reads the
per-dispatch ``*kernel_trace.csv``.
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


def _grid(r) -> str:
    for k in ("Grid_Size", "Grid_Size_X"):
        if k in r:
            return "x".join(r[c] for c in ("Grid_Size_X", "Grid_Size_Y", "Grid_Size_Z") if c in r) \
                if k == "Grid_Size_X" else r[k]
    return "?"


def stats_by_grid(d: str, out=sys.stdout, top: int = 24):
    t = _find(d, "*kernel_trace.csv")
    if t is None:
        raise SystemExit(f"no kernel_trace.csv under {d}")
    agg = collections.defaultdict(list)
    with open(t) as fh:
        for r in csv.DictReader(fh):
            agg[(r["Kernel_Name"][:72], _grid(r))].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    rows = sorted(agg.items(), key=lambda kv: -sum(kv[1]))
    total = sum(sum(v) for _, v in rows)
    for (k, g), v in rows[:top]:
        print(f"{k:72s} grid {g} calls {len(v)} total_ms {sum(v) / 1e6:.3f} avg_us {sum(v) / len(v) / 1e3:.2f} "
              f"min_us {min(v) / 1e3:.2f} max_us {max(v) / 1e3:.2f} pct {100.0 * sum(v) / total:.2f}", file=out)
    return t


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


def pmc(d: str, out=sys.stdout, by_grid: bool = False):
    f = _find(d, "*counter_collection.csv")
    if f is None:
        raise SystemExit(f"no counter_collection.csv under {d}")
    agg = collections.defaultdict(list)
    with open(f) as fh:
        for r in csv.DictReader(fh):
            name = r["Kernel_Name"][:72] + (f" grid {_grid(r)}" if by_grid else "")
            agg[(name, r["Counter_Name"])].append(float(r["Counter_Value"]))
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
    ap.add_argument("--by-grid", action="store_true", help="one line per kernel and grid size")
    a = ap.parse_args(argv)
    if a.mode == "stats":
        big = stats_by_grid(a.dir) if a.by_grid else stats(a.dir)
    else:
        big = pmc(a.dir, by_grid=a.by_grid)
    if a.prune:
        for f in ([big] if big else []) + glob.glob(os.path.join(a.dir, "**", "*.db"), recursive=True):
            os.remove(f)
    return 0


if __name__ == "__main__":
    sys.exit(main())
