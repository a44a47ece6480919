"""Summaries of rocprofv3 CSV output as markdown tables.

    python -m citcom_amd.prof_summary stats  DIR/NAME_kernel_stats.csv [--top 25]
    python -m citcom_amd.prof_summary trace  DIR/NAME_kernel_trace.csv [--regex k_gs_colour]
    python -m citcom_amd.prof_summary counters DIR/NAME_counter_collection.csv [--regex k_gs_colour]

``stats``: the per-kernel table of ``rocprofv3 --kernel-trace --stats``
(calls, total and average duration, share of kernel time).  ``trace``: the
kernel trace grouped by kernel name and grid size (one multigrid level's
launches of a kernel share a grid size), with the launches' count and
average and total duration.  ``counters``: a ``--pmc`` run grouped the same
way, with the average counter value per dispatch (FETCH_SIZE and WRITE_SIZE
are in KiB).  Only the standard library is used, so the summaries can be
made where the profile was taken and the large CSV files left there.
"""
from __future__ import annotations

import argparse
import csv
import re
import sys
from collections import defaultdict


def _col(row: dict, *names: str):
    for n in names:
        if n in row:
            return row[n]
    raise KeyError(f"none of {names} in the CSV header {list(row)}")


def _short(name: str, width: int = 70) -> str:
    name = name.replace("void ", "").replace("citcom_hip::", "")
    name = re.sub(r"\(.*\)$", "", name).strip()
    return name if len(name) <= width else name[:width - 3] + "..."


def _grid(row: dict) -> int:
    if "Grid_Size" in row:
        return int(row["Grid_Size"])
    return int(_col(row, "Grid_Size_X")) * int(row.get("Grid_Size_Y", 1) or 1) * int(row.get("Grid_Size_Z", 1) or 1)


def stats(path: str, top: int) -> str:
    rows = list(csv.DictReader(open(path)))
    total = sum(float(_col(r, "TotalDurationNs")) for r in rows)
    rows.sort(key=lambda r: -float(_col(r, "TotalDurationNs")))
    out = ["| kernel | calls | total ms | average µs | share |", "|---|---:|---:|---:|---:|"]
    for r in rows[:top]:
        t = float(_col(r, "TotalDurationNs"))
        out.append(f"| `{_short(_col(r, 'Name', 'Kernel_Name'))}` | {int(_col(r, 'Calls'))} | {t / 1e6:.2f} | "
                   f"{float(_col(r, 'AverageNs')) / 1e3:.2f} | {100 * t / total:.1f} % |")
    out.append(f"\nAll kernels: {total / 1e6:.1f} ms in {sum(int(_col(r, 'Calls')) for r in rows)} launches.")
    return "\n".join(out)


def trace(path: str, regex: str) -> str:
    pat = re.compile(regex)
    acc = defaultdict(lambda: [0, 0.0])
    for r in csv.DictReader(open(path)):
        name = _col(r, "Kernel_Name")
        if not pat.search(name):
            continue
        d = float(_col(r, "End_Timestamp")) - float(_col(r, "Start_Timestamp"))
        a = acc[(_short(name), _grid(r))]
        a[0] += 1
        a[1] += d
    out = ["| kernel | grid (work-items) | launches | average µs | total ms |", "|---|---:|---:|---:|---:|"]
    for (name, grid), (n, t) in sorted(acc.items(), key=lambda kv: (kv[0][0], -kv[0][1])):
        out.append(f"| `{name}` | {grid} | {n} | {t / n / 1e3:.2f} | {t / 1e6:.2f} |")
    return "\n".join(out)


def counters(path: str, regex: str) -> str:
    pat = re.compile(regex)
    acc = defaultdict(lambda: [0, 0.0])
    for r in csv.DictReader(open(path)):
        name = _col(r, "Kernel_Name")
        if not pat.search(name):
            continue
        a = acc[(_short(name), _grid(r), _col(r, "Counter_Name"))]
        a[0] += 1
        a[1] += float(_col(r, "Counter_Value"))
    out = ["| kernel | grid (work-items) | counter | dispatches | average per dispatch |", "|---|---:|---|---:|---:|"]
    for (name, grid, ctr), (n, v) in sorted(acc.items(), key=lambda kv: (kv[0][0], -kv[0][1], kv[0][2])):
        out.append(f"| `{name}` | {grid} | {ctr} | {n} | {v / n:.1f} |")
    return "\n".join(out)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("kind", choices=("stats", "trace", "counters"))
    ap.add_argument("csv")
    ap.add_argument("--top", type=int, default=25)
    ap.add_argument("--regex", default=".")
    a = ap.parse_args(argv)
    if a.kind == "stats":
        print(stats(a.csv, a.top))
    elif a.kind == "trace":
        print(trace(a.csv, a.regex))
    else:
        print(counters(a.csv, a.regex))
    return 0


if __name__ == "__main__":
    sys.exit(main())
