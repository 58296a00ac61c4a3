"""Per-kernel launch statistics from a rocprofv3 kernel trace (--kernel-trace --stats).

    python stable-fluids_amd/trace_record.py run_kernel_stats.csv --cells 16777216 \
        --workload "plume 256^3, D=40 P=80 RBGS" --json OUT.json --md OUT.md --command "..."

Reads the run_kernel_stats.csv of one run and lists, per kernel, the calls, the mean,
minimum and maximum duration and, for the kernels with known algorithmic bytes per cell
(traffic_record.ALGORITHMIC_BYTES_PER_CELL and the black-cell passes), the mean
algorithmic bandwidth and its fraction of the 8 TB/s HBM peak.
"""
import argparse
import csv
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import traffic_record  # noqa: E402

HBM_PEAK_GBS = 8000.0


def build(stats_csv, cells, workload):
    kernels = {}
    for row in csv.DictReader(open(stats_csv)):
        name = traffic_record.short_name(row["Name"])
        rec = {"calls": int(row["Calls"]), "total_ms": round(float(row["TotalDurationNs"]) / 1e6, 3),
               "avg_us": round(float(row["AverageNs"]) / 1e3, 3), "min_us": round(float(row["MinNs"]) / 1e3, 3),
               "max_us": round(float(row["MaxNs"]) / 1e3, 3), "percent": round(float(row["Percentage"]), 2)}
        alg = traffic_record.algorithmic_bytes(name, cells)
        if alg is not None:
            gbs = alg / (float(row["AverageNs"]) * 1e-9) / 1e9
            rec.update({"algorithmic_bytes": alg, "gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)})
        kernels[name] = rec
    return {"workload": workload, "cells": cells, "kernels": kernels}


def markdown(rec, commands, top):
    lines = [f"# Kernel trace: {rec['workload']}", "", "Command (from the repository root on one MI355X):", "", "```",
             *commands, "```", "",
             "| kernel | calls | total ms | avg us | min us | max us | % | algorithmic MB | GB/s | of 8 TB/s |",
             "|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|"]
    items = sorted(rec["kernels"].items(), key=lambda kv: -kv[1]["total_ms"])[:top]
    for name, k in items:
        alg = f"{k['algorithmic_bytes'] / 1e6:.1f}" if "algorithmic_bytes" in k else ""
        gbs = f"{k['gbs']:.0f}" if "gbs" in k else ""
        frac = f"{k['frac']:.3f}" if "frac" in k else ""
        lines.append(f"| `{name}` | {k['calls']} | {k['total_ms']:.2f} | {k['avg_us']:.2f} | {k['min_us']:.2f} | "
                     f"{k['max_us']:.2f} | {k['percent']:.1f} | {alg} | {gbs} | {frac} |")
    return "\n".join(lines) + "\n"


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("stats_csv")
    ap.add_argument("--cells", type=int, required=True)
    ap.add_argument("--workload", default="")
    ap.add_argument("--json")
    ap.add_argument("--md")
    ap.add_argument("--top", type=int, default=30)
    ap.add_argument("--command", action="append", default=[], help="command line to record in the summary")
    a = ap.parse_args()
    rec = build(a.stats_csv, a.cells, a.workload)
    if a.json:
        with open(a.json, "w") as f:
            json.dump(rec, f, indent=1)
    if a.md:
        with open(a.md, "w") as f:
            f.write(markdown(rec, a.command, a.top))
    if not a.json and not a.md:
        print(json.dumps(rec, indent=1))


if __name__ == "__main__":
    main()
