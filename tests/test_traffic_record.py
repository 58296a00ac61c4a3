"""The HBM traffic record: counter calibration on the copy kernel and algorithmic bytes."""
import csv
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
spec = importlib.util.spec_from_file_location("traffic_record", os.path.join(ROOT, "stable-fluids_amd", "traffic_record.py"))
tr = importlib.util.module_from_spec(spec)
spec.loader.exec_module(tr)

FIELDS = ["Kernel_Name", "Counter_Name", "Counter_Value"]
CELLS = 1 << 20


def write_csv(path, counter, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=FIELDS)
        w.writeheader()
        for name, value in rows:
            w.writerow({"Kernel_Name": f"void sf::(anonymous namespace)::{name}(float*, unsigned long)", "Counter_Name": counter,
                        "Counter_Value": value})


def test_calibration_and_bytes(tmp_path):
    kib = 4 * CELLS / 1024
    # the copy kernel counts half of its fetched bytes, all of its written bytes
    write_csv(tmp_path / "f.csv", "FETCH_SIZE", [("k_copy_scaled", kib / 2), ("k_copy_scaled", kib / 2),
                                                 ("k_rbgs_split_pass<0, false, sf::StoreF32, true>", kib),
                                                 ("k_rbgs_split_pass<1, true, sf::StoreF32, true>", kib / 2),
                                                 ("k_unlisted", 1.0)])
    write_csv(tmp_path / "w.csv", "WRITE_SIZE", [("k_copy_scaled", kib), ("k_copy_scaled", kib),
                                                 ("k_rbgs_split_pass<0, false, sf::StoreF32, true>", kib),
                                                 ("k_rbgs_split_pass<1, true, sf::StoreF32, true>", kib),
                                                 ("k_unlisted", 1.0)])
    rec = tr.build(str(tmp_path / "f.csv"), str(tmp_path / "w.csv"), CELLS, "test")
    assert rec["fetch_factor"] == pytest.approx(2.0) and rec["write_factor"] == pytest.approx(1.0)
    k = rec["kernels"]
    assert "k_unlisted" not in k
    assert k["k_copy_scaled"]["dispatches"] == 2
    assert k["k_copy_scaled"]["traffic_bytes"] == k["k_copy_scaled"]["algorithmic_bytes"] == 8 * CELLS
    # two fields fetched (x_old, b) and one written: the algorithmic minimum exactly
    assert k["k_rbgs_split_pass<0, false, sf::StoreF32, true>"]["traffic_bytes"] == 12 * CELLS
    # first pass from a zero guess reads only b
    assert k["k_rbgs_split_pass<1, true, sf::StoreF32, true>"]["algorithmic_bytes"] == 8 * CELLS
    md = tr.markdown(rec, ["cmd"])
    assert "| `k_copy_scaled` | 2 |" in md


def test_committed_record_has_the_bench_kernel():
    import json
    rec = json.load(open(os.path.join(ROOT, "profiles", "traffic_plume256_kernels.json")))
    src = open(os.path.join(ROOT, "bench.py")).read()
    key = src.split('VELOCITY_PASS_KERNEL = "')[1].split('"')[0]
    assert key in rec["kernels"]
    assert rec["kernels"][key]["traffic_bytes"] >= rec["kernels"][key]["algorithmic_bytes"]


def test_black_cell_pass_bytes():
    """b (4 B) + the iterate read (zero guess: none; natural: 4 B; black cells: 2 B) + the
    iterate written (black cells: 2 B; natural: 4 B)."""
    for args, per in (("1, 0, 0", 6), ("0, 2, 0", 8), ("0, 1, 0", 10), ("1, 2, 1", 10), ("0, 1, 1", 12)):
        assert tr.algorithmic_bytes(f"k_rbgs_half<{args}, 64, 8, 2>", CELLS) == per * CELLS, args


def test_trace_record(tmp_path):
    spec2 = importlib.util.spec_from_file_location("trace_record", os.path.join(ROOT, "stable-fluids_amd", "trace_record.py"))
    trr = importlib.util.module_from_spec(spec2)
    spec2.loader.exec_module(trr)
    path = tmp_path / "stats.csv"
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage", "MinNs", "MaxNs", "StdDev"])
        w.writeheader()
        w.writerow({"Name": "void sf::(anonymous namespace)::k_rbgs_half<0, 2, 0, 64, 8, 2>(sf::HalfPassArgs)", "Calls": 10,
                    "TotalDurationNs": 10 * 1048.576, "AverageNs": 1048.576, "Percentage": 90.0, "MinNs": 1000, "MaxNs": 1100, "StdDev": 1})
        w.writerow({"Name": "k_other(int)", "Calls": 1, "TotalDurationNs": 5, "AverageNs": 5, "Percentage": 10.0, "MinNs": 5, "MaxNs": 5, "StdDev": 0})
    rec = trr.build(str(path), CELLS, "test")
    k = rec["kernels"]["k_rbgs_half<0, 2, 0, 64, 8, 2>"]
    # 8 B per cell of 2^20 cells in 1.048576 us: 8 TB/s
    assert k["calls"] == 10 and k["algorithmic_bytes"] == 8 * CELLS and k["frac"] == pytest.approx(1.0)
    assert "gbs" not in rec["kernels"]["k_other"]
    assert "| `k_rbgs_half<0, 2, 0, 64, 8, 2>` | 10 |" in trr.markdown(rec, ["cmd"], 5)
