"""HBM traffic per kernel launch from two rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE).

    python stable-fluids_amd/traffic_record.py FETCH.csv WRITE.csv --cells 16777216 \
        --workload "plume 256^3, D=40 P=80 RBGS" --json OUT.json --md OUT.md

Each CSV is the run_counter_collection.csv of one `rocprofv3 --pmc <counter>` pass over
the same program. Counter values are KiB per dispatch. Both are calibrated on
k_copy_scaled, which reads one field and writes one field (4 B per cell each): fetched
bytes = FETCH_SIZE * 1024 * (4 B * cells) / (1024 * FETCH_SIZE of the copy), likewise for
writes.
The algorithmic bytes per cell of every stage kernel (each input field read once, each
output field written once) are listed in ALGORITHMIC_BYTES_PER_CELL.
"""
import argparse
import collections
import csv
import json

# bytes per cell a launch must move at minimum (4-byte fields)
ALGORITHMIC_BYTES_PER_CELL = {
    "k_rbgs_split_pass": 12,      # read x_old, b; write x_new
    "k_jacobi_pass": 12,          # read x_old, b; write x_new
    "k_copy_scaled": 8,           # read src; write dst
    "k_advect_velocity": 24,      # read 3 velocity components; write 3
    "k_advect_scalar": 20,        # read the scalar and 3 velocity components; write 1
    "k_advect_velocity_b": 24,    # buffer-resource variants of the two above
    "k_advect_scalar_b": 20,
    "k_project": 36,              # read a (3 components), p; write v (3), div, |v|
    "k_divergence": 16,           # read 3 velocity components; write div
    "k_project_z": 36,            # z-marching variants of the two above
    "k_divergence_z": 16,
    "k_add_force": 36,            # read 3 velocity + 3 force components; write 3
    "k_add_scaled": 12,           # read the scalar and its source; write 1
}
# first pass from a zero guess: x_old is not read
ZERO_GUESS_BYTES_PER_CELL = 8
# black-cell RBGS passes k_rbgs_half<solve, in, out, ...>: b (4 B per cell) plus the
# input iterate (in 0: zero guess, none; 1: the natural field, 4 B; 2: the black cells,
# 2 B) plus the output (out 0: the black cells, 2 B; 1: the natural field, 4 B)
HALF_IN_BYTES = {0: 0, 1: 4, 2: 2}
HALF_OUT_BYTES = {0: 2, 1: 4}


def half_pass_bytes_per_cell(name):
    args = [a.strip() for a in name.split("<")[1].split(">")[0].split(",")]
    return 4 + HALF_IN_BYTES[int(args[1])] + HALF_OUT_BYTES[int(args[2])]


def short_name(kernel_name):
    return kernel_name.replace("sf::(anonymous namespace)::", "").replace("void ", "").split("(")[0]


def algorithmic_bytes(name, cells):
    base = name.split("<")[0]
    if base == "k_rbgs_half":
        return half_pass_bytes_per_cell(name) * cells
    per = ALGORITHMIC_BYTES_PER_CELL.get(base)
    if per is None:
        return None
    if base in ("k_rbgs_split_pass", "k_jacobi_pass") and name.split("<")[1].split(",")[1].strip() == "true":
        per = ZERO_GUESS_BYTES_PER_CELL
    return per * cells


def mean_counter(path, counter):
    vals = collections.defaultdict(list)
    for row in csv.DictReader(open(path)):
        if row["Counter_Name"] == counter:
            vals[short_name(row["Kernel_Name"])].append(float(row["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in vals.items()}, {k: len(v) for k, v in vals.items()}


def build(fetch_csv, write_csv, cells, workload):
    fetch, n_fetch = mean_counter(fetch_csv, "FETCH_SIZE")
    write, _ = mean_counter(write_csv, "WRITE_SIZE")
    if "k_copy_scaled" not in fetch or "k_copy_scaled" not in write:
        raise SystemExit("k_copy_scaled is missing from the counter runs: no calibration")
    field = 4.0 * cells
    ff = field / (fetch["k_copy_scaled"] * 1024.0)
    wf = field / (write["k_copy_scaled"] * 1024.0)
    kernels = {}
    for name in fetch:
        alg = algorithmic_bytes(name, cells)
        if alg is None or name not in write:
            continue
        t = fetch[name] * 1024.0 * ff + write[name] * 1024.0 * wf
        kernels[name] = {"dispatches": n_fetch[name], "fetch_kib": round(fetch[name], 1), "write_kib": round(write[name], 1),
                         "traffic_bytes": int(round(t)), "algorithmic_bytes": alg}
    return {"workload": workload, "cells": cells, "fetch_factor": ff, "write_factor": wf, "kernels": kernels}


def markdown(rec, commands):
    lines = [f"# HBM traffic per launch: {rec['workload']} (FETCH_SIZE / WRITE_SIZE)", "",
             "Commands (from the repository root on one MI355X; one counter per pass):", "", "```", *commands, "```", "",
             f"Calibration on `k_copy_scaled` (reads and writes one {4 * rec['cells'] / 2**20:.0f} MiB field): "
             f"fetched bytes = FETCH_SIZE x 1024 x {rec['fetch_factor']:.4f}, written bytes = WRITE_SIZE x 1024 x "
             f"{rec['write_factor']:.4f}.", "",
             "| kernel | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | traffic MB (calibrated) | algorithmic MB | ratio |",
             "|---|---:|---:|---:|---:|---:|---:|"]
    for name, k in rec["kernels"].items():
        lines.append(f"| `{name}` | {k['dispatches']} | {k['fetch_kib']:.0f} | {k['write_kib']:.0f} | "
                     f"{k['traffic_bytes'] / 1e6:.1f} | {k['algorithmic_bytes'] / 1e6:.1f} | "
                     f"{k['traffic_bytes'] / k['algorithmic_bytes']:.2f} |")
    return "\n".join(lines) + "\n"


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("fetch_csv")
    ap.add_argument("write_csv")
    ap.add_argument("--cells", type=int, required=True)
    ap.add_argument("--workload", default="")
    ap.add_argument("--json")
    ap.add_argument("--md")
    ap.add_argument("--command", action="append", default=[], help="command line to record in the summary")
    a = ap.parse_args()
    rec = build(a.fetch_csv, a.write_csv, a.cells, a.workload)
    if a.json:
        with open(a.json, "w") as f:
            json.dump(rec, f, indent=1)
    if a.md:
        with open(a.md, "w") as f:
            f.write(markdown(rec, a.command))
    if not a.json and not a.md:
        print(json.dumps(rec, indent=1))


if __name__ == "__main__":
    main()
