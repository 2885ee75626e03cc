"""The rocprofv3 CSV summaries behind the committed profiles: kernel stats,
HBM traffic scaling, counter-derived figures and trace windows, on small
synthetic CSVs whose answers are known."""
import csv
import os
import sys

import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "dagger.jl_amd"))
from dagger_mi355x import rocprof_summary as R  # noqa: E402

CC_FIELDS = ["Correlation_Id", "Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value",
             "Start_Timestamp", "End_Timestamp"]


def _write(path, fields, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=fields)
        w.writeheader()
        for r in rows:
            w.writerow(r)


def test_stats_orders_by_total_time(tmp_path):
    p = tmp_path / "s.csv"
    _write(p, ["Name", "Calls", "TotalDurationNs", "AverageNs"],
           [{"Name": "a", "Calls": 2, "TotalDurationNs": 1e6, "AverageNs": 5e5},
            {"Name": "b", "Calls": 1, "TotalDurationNs": 3e6, "AverageNs": 3e6}])
    out = R.stats(str(p), 5).splitlines()
    assert out[1].split()[-1] == "b" and out[2].split()[-1] == "a"
    assert "4.000 ms over 3 dispatches" in out[-1]


def test_traffic_scales_by_the_copy_kernel(tmp_path):
    f, w = tmp_path / "f.csv", tmp_path / "w.csv"
    # copy of 1 MiB: FETCH reads half its bytes, WRITE all of them (KiB)
    _write(f, CC_FIELDS, [
        dict(Correlation_Id=1, Dispatch_Id=1, Kernel_Name="__amd_rocclr_copyBuffer",
             Counter_Name="FETCH_SIZE", Counter_Value=512, Start_Timestamp=0, End_Timestamp=1),
        dict(Correlation_Id=2, Dispatch_Id=2, Kernel_Name="gemm_kernel<x>",
             Counter_Name="FETCH_SIZE", Counter_Value=100, Start_Timestamp=0, End_Timestamp=1)])
    _write(w, CC_FIELDS, [
        dict(Correlation_Id=1, Dispatch_Id=1, Kernel_Name="__amd_rocclr_copyBuffer",
             Counter_Name="WRITE_SIZE", Counter_Value=1024, Start_Timestamp=0, End_Timestamp=1),
        dict(Correlation_Id=2, Dispatch_Id=2, Kernel_Name="gemm_kernel<x>",
             Counter_Name="WRITE_SIZE", Counter_Value=10, Start_Timestamp=0, End_Timestamp=1)])
    out = R.traffic(str(f), str(w), "gemm_kernel", 1 << 20)
    assert "fetch_scale: 2.0000" in out and "write_scale: 1.0000" in out
    assert f"traffic_bytes_per_launch: {200 * 1024 + 10 * 1024}" in out


def test_counters_derive_clock_and_mfma_utilisation(tmp_path):
    p = tmp_path / "c.csv"
    dur = 1_000_000  # 1 ms
    rows = []
    # GRBM_GUI_ACTIVE summed over 8 XCDs: 8 x 2.0e6 cycles in 1 ms = 2000 MHz
    for c, v in (("GRBM_GUI_ACTIVE", 8 * 2.0e6), ("SQ_VALU_MFMA_BUSY_CYCLES", 0.5 * 2.0e6 * 1024),
                 ("SQ_LDS_BANK_CONFLICT", 25.0), ("SQ_LDS_IDX_ACTIVE", 100.0),
                 ("SQ_WAVE_CYCLES", 1000.0), ("SQ_WAIT_ANY", 250.0)):
        rows.append(dict(Correlation_Id=1, Dispatch_Id=7, Kernel_Name="gemm_kernel<d>",
                         Counter_Name=c, Counter_Value=v, Start_Timestamp=0, End_Timestamp=dur))
    _write(p, CC_FIELDS, rows)
    out = R.counters([str(p)], "gemm_kernel", 1024, 8)
    assert "mfma_util: 0.5000" in out
    assert "avg_clock_mhz: 2000" in out
    assert "lds_bank_conflict_share: 0.2500" in out
    assert "sq_wait_any_per_wave_cycle: 0.2500" in out


def test_window_splits_a_trace(tmp_path):
    p = tmp_path / "t.csv"
    names = ["potrf_x", "trtri_leaf_kernel", "gemm_a", "gemm_b", "potrf_x", "trtri_leaf_kernel",
             "gemm_a"]
    rows = [dict(Kernel_Name=n, Start_Timestamp=1000 * i, End_Timestamp=1000 * i + 500)
            for i, n in enumerate(names)]
    _write(p, ["Kernel_Name", "Start_Timestamp", "End_Timestamp"], rows)
    out = R.window(str(p), "trtri_leaf", "potrf_").splitlines()
    assert out[0] == "window 0: span 0.003 ms, 3 dispatches"
    assert out[1] == "window 1: span 0.002 ms, 2 dispatches"
    assert out[2] == "window 1, per kernel:"
    first = R.window(str(p), "trtri_leaf", "potrf_", index=0).splitlines()
    assert first[2] == "window 0, per kernel:"
    assert any("gemm_b" in line for line in first[3:])
    assert not any("gemm_b" in line for line in out[3:])
    with pytest.raises(SystemExit):
        R.window(str(p), "nothing_matches", "potrf_")
