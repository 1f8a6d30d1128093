"""CPU tests of bench.py's measurement helpers: the rocprofv3 CSV parsing
behind roofline.traffic and roofline.valu_issue, on small files written
here in rocprofv3's column layout (no GPU, no profiler)."""
import csv
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

FWD = "void (anonymous namespace)::k_forward<32, 3, false, false, false, true, true>(gt_dev_scene, ...)"


def _write_counters(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for r in rows:
            w.writerow(r)


def test_per_dispatch_sums_rows_of_one_dispatch(tmp_path):
    """A counter read per XCD or per shader engine comes as several rows of
    one dispatch: they are summed, dispatches stay in dispatch order, other
    counters are ignored."""
    d = tmp_path / "run"
    d.mkdir()
    _write_counters(d / "run_counter_collection.csv", [
        {"Dispatch_Id": 7, "Kernel_Name": FWD, "Counter_Name": "FETCH_SIZE", "Counter_Value": 10.0},
        {"Dispatch_Id": 3, "Kernel_Name": "k_colors", "Counter_Name": "FETCH_SIZE", "Counter_Value": 1.5},
        {"Dispatch_Id": 7, "Kernel_Name": FWD, "Counter_Name": "FETCH_SIZE", "Counter_Value": 2.5},
        {"Dispatch_Id": 7, "Kernel_Name": FWD, "Counter_Name": "WRITE_SIZE", "Counter_Value": 99.0},
    ])
    rows = bench._counter_csv_rows(str(tmp_path))
    assert bench._per_dispatch(rows, "FETCH_SIZE") == [("k_colors", 1.5), (FWD, 12.5)]
    assert bench._per_dispatch(rows, "WRITE_SIZE") == [(FWD, 99.0)]
    assert bench._per_dispatch(rows, "SQ_WAVES") == []


def test_agent_info_reads_the_gpu_agent(tmp_path):
    with open(tmp_path / "run_agent_info.csv", "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Node_Id", "Simd_Count", "Num_Xcc", "Name"])
        w.writeheader()
        w.writerow({"Node_Id": 0, "Simd_Count": 0, "Num_Xcc": 0, "Name": "cpu"})
        w.writerow({"Node_Id": 1, "Simd_Count": 1024, "Num_Xcc": 8, "Name": "gfx950"})
    assert bench._agent_info(str(tmp_path)) == (1024, 8)
    assert bench._agent_info(str(tmp_path / "missing")) is None


def test_valu_issue_figures():
    """busy_frac = 4 x active VALU quad-cycles / (SIMDs x cycles per XCD);
    lane efficiency = thread-cycles / (64 x active quad-cycles); the clock is
    the per-XCD cycles over the traced duration.  Only the first two
    k_forward dispatches count (the probe's plain launches)."""
    simds, xcds = 1024, 8
    cycles = 20_000_000.0  # per XCD
    active = 0.75 * simds * cycles / 4.0  # 75 % busy
    valu = {
        "SQ_INSTS_VALU": [(FWD, 5e9), ("k_colors", 1.0), (FWD, 5e9), (FWD, 9e9)],
        "SQ_ACTIVE_INST_VALU": [(FWD, active), (FWD, active), (FWD, 1.0)],
        "SQ_THREAD_CYCLES_VALU": [(FWD, 0.5 * 64 * active), (FWD, 0.5 * 64 * active)],
        "GRBM_GUI_ACTIVE": [(FWD, cycles * xcds), (FWD, cycles * xcds)],
    }
    out = bench._valu_issue(valu, (simds, xcds), [10.0, 10.0])
    assert out["valu_insts"] == 5_000_000_000
    assert out["busy_frac"] == pytest.approx(0.75)
    assert out["lane_efficiency"] == pytest.approx(0.5)
    assert out["cycles_per_xcd"] == int(cycles)
    assert out["clock_ghz"] == pytest.approx(2.0)
    assert bench._valu_issue(valu, None, [10.0]) is None
    short = dict(valu, SQ_INSTS_VALU=[(FWD, 5e9)])
    assert bench._valu_issue(short, (simds, xcds), [10.0, 10.0]) is None
