"""Summarise rocprofv3 output (kernel trace and PMC counters) as text tables.

Usage::

    python -m prince_amd.rocprof_summary <rocprofv3 output dir> [--pmc COUNTER]
    python -m prince_amd.rocprof_summary traffic --fetch DIR --write DIR --match NAME ...

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


def traffic(fetch_dir, write_dir, match, calib_bytes, calib_match="elementwise"):
    """Per-dispatch HBM traffic of the kernels whose names contain one of
    ``match``, from a FETCH_SIZE run and a WRITE_SIZE run (KiB per dispatch).

    The calibration factors are FETCH_SIZE and WRITE_SIZE of the largest
    kernel whose name contains ``calib_match`` (an elementwise copy that reads
    and writes ``calib_bytes``) divided by ``calib_bytes``; traffic in bytes is
    FETCH / fetch_factor + WRITE / write_factor.
    """
    fetch = counter_rows(fetch_dir, "FETCH_SIZE")
    write = counter_rows(write_dir, "WRITE_SIZE")
    cal = [(max(v), k) for k, v in fetch.items() if calib_match in k]
    if not cal:
        raise SystemExit("no calibration kernel in the FETCH_SIZE run")
    _, cname = max(cal)
    cf = max(fetch[cname]) * 1024.0
    cw = max(write.get(cname, [0.0])) * 1024.0
    out = {"calibration_kernel": _short(cname, 200), "calibration_bytes_each_way": calib_bytes,
           "calibration_fetch_KiB": cf / 1024.0, "calibration_write_KiB": cw / 1024.0,
           "fetch_factor": cf / calib_bytes, "write_factor": cw / calib_bytes, "kernels": {}}
    for name, vals in fetch.items():
        short = _short(name, 200)
        if not any(m in short for m in match):
            continue
        wv = write.get(name, [])
        out["kernels"][short] = {
            "dispatches": len(vals),
            "FETCH_SIZE_KiB": sum(vals) / len(vals),
            "WRITE_SIZE_KiB": (sum(wv) / len(wv)) if wv else 0.0,
        }
    return out


def main():
    import json
    import sys

    if len(sys.argv) > 1 and sys.argv[1] == "traffic":
        ap = argparse.ArgumentParser(prog="rocprof_summary traffic")
        ap.add_argument("--fetch", required=True)
        ap.add_argument("--write", required=True)
        ap.add_argument("--match", action="append", required=True)
        ap.add_argument("--calib-bytes", type=float, default=256 * 1024 * 1024)
        a = ap.parse_args(sys.argv[2:])
        print(json.dumps(traffic(a.fetch, a.write, a.match, a.calib_bytes), indent=1))
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("path")
    ap.add_argument("--pmc", action="append", default=[])
    a = ap.parse_args()
    print(summarize(a.path, a.pmc))


if __name__ == "__main__":
    main()
