"""Summarise rocprofv3 output (kernel trace and PMC counters) as text tables.

Usage::

    python -m prince_amd.rocprof_summary <rocprofv3 output dir> [--pmc COUNTER]

Reads every ``*_results.db`` (rocpd SQLite) or ``*kernel_trace.csv`` /
``*counter_collection.csv`` under the directory and prints, per kernel: the
dispatch count, average / min / max duration and total time, and -- when PMC
records are present -- the average counter value per dispatch.
"""

from __future__ import annotations

import argparse
import csv
import glob
import os
import sqlite3
from collections import defaultdict


def _short(name, width=90):
    name = name.replace("(anonymous namespace)::", "")
    return name if len(name) <= width else name[: width - 3] + "..."


def kernel_rows(path):
    """[(name, start_ns, end_ns)] from a rocprofv3 output directory."""
    rows = []
    for db in glob.glob(os.path.join(path, "**", "*_results.db"), recursive=True):
        con = sqlite3.connect(db)
        try:
            for name, start, end in con.execute("select name, start, end from kernels"):
                rows.append((name, int(start), int(end)))
        finally:
            con.close()
    for fn in glob.glob(os.path.join(path, "**", "*kernel_trace.csv"), recursive=True):
        with open(fn) as f:
            for r in csv.DictReader(f):
                rows.append((r["Kernel_Name"], int(r["Start_Timestamp"]), int(r["End_Timestamp"])))
    return rows


def counter_rows(path, counter):
    """{kernel name: [value per dispatch]} for one PMC counter."""
    out = defaultdict(list)
    for fn in glob.glob(os.path.join(path, "**", "*counter_collection.csv"), recursive=True):
        with open(fn) as f:
            for r in csv.DictReader(f):
                if r.get("Counter_Name") == counter:
                    out[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    for db in glob.glob(os.path.join(path, "**", "*_results.db"), recursive=True):
        con = sqlite3.connect(db)
        try:
            tabs = {t for (t,) in con.execute("select name from sqlite_master")}
            if "counters_collection" in tabs:
                q = ("select kernel_name, counter_value from counters_collection "
                     "where counter_name = ?")
                for name, val in con.execute(q, (counter,)):
                    out[name].append(float(val))
        finally:
            con.close()
    return out


def summarize(path, counters=()):
    rows = kernel_rows(path)
    agg = defaultdict(list)
    for name, s, e in rows:
        agg[name].append(e - s)
    lines = [f"{'kernel':90s} {'calls':>7s} {'avg_us':>10s} {'min_us':>9s} {'max_us':>9s} {'total_ms':>10s}"]
    for name, ds in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        lines.append(
            f"{_short(name):90s} {len(ds):7d} {sum(ds)/len(ds)/1e3:10.2f} {min(ds)/1e3:9.2f} "
            f"{max(ds)/1e3:9.2f} {sum(ds)/1e6:10.3f}"
        )
    for c in counters:
        vals = counter_rows(path, c)
        if not vals:
            continue
        lines.append("")
        lines.append(f"{'kernel':90s} {'dispatches':>10s} {c + ' avg':>18s}")
        for name, v in sorted(vals.items(), key=lambda kv: -sum(kv[1])):
            lines.append(f"{_short(name):90s} {len(v):10d} {sum(v)/len(v):18.1f}")
    return "\n".join(lines)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("path")
    ap.add_argument("--pmc", action="append", default=[])
    a = ap.parse_args()
    print(summarize(a.path, a.pmc))


if __name__ == "__main__":
    main()
