"""The reductions of rocprofv3 output kept under profiles/ (forge_amd.counter_summary,
forge_amd.trace_summary), on synthetic inputs whose answers are known: per-kernel
totals, the copy-kernel calibration of FETCH_SIZE / WRITE_SIZE, the matrix-core
busy fraction, and the k_ts_q2 rate from a kernel trace."""
import csv
import json
import sqlite3
import sys

import pytest

from forge_amd import counter_summary as cs
from forge_amd import trace_summary as ts


def _write_csv(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=list(rows[0]))
        w.writeheader()
        w.writerows(rows)


def test_kernels_from_stats_csv_and_db(tmp_path):
    stats = tmp_path / "run_kernel_stats.csv"
    _write_csv(stats, [
        {"Name": "void forge::k_a(long)", "Calls": 4, "TotalDurationNs": 4000},
        {"Name": "forge::k_b<HIP_vector_type<float, 2u> >(int)", "Calls": 1, "TotalDurationNs": 12000},
    ])
    lines = cs.kernels(str(stats))
    assert lines[0] == "total kernel time 0.0 ms, 5 dispatches"
    # sorted by total time; float2 spelled out; mean per call in us
    assert "k_b<float2 >" in lines[1] and "75.0%" in lines[1]
    assert lines[2].split()[2] == "4" and "1.0 us" in lines[2]

    db = tmp_path / "run_results.db"
    con = sqlite3.connect(db)
    con.execute("create table kernels (name text, start integer, end integer)")
    con.executemany("insert into kernels values (?, ?, ?)",
                    [("void forge::k_a(long)", 0, 1000), ("void forge::k_a(long)", 5000, 8000)])
    con.commit()
    con.close()
    lines = cs.kernels(str(db))
    assert lines[0] == "total kernel time 0.0 ms, 2 dispatches"
    assert "2.0 us" in lines[1] and "100.0%" in lines[1]


def _counter_rows(kernel, counter, per_dispatch, ndisp, dur_ns, first_id=1):
    return [{"Kernel_Name": kernel, "Counter_Name": counter, "Counter_Value": per_dispatch,
             "Dispatch_Id": first_id + q, "Start_Timestamp": 0, "End_Timestamp": dur_ns}
            for q in range(ndisp)]


def test_traffic_calibrated_on_the_copy_kernel(tmp_path):
    n, K, eb = 100, 2, 8
    m = (n + 2 * (min(20, n) // 2)) // 2  # nUpper
    copy_bytes = n * m * 8 * K  # k_bt_promote reads the real eigenvectors once
    # counters read 2x the true reads and exactly the writes
    fetch = (_counter_rows("forge::k_bt_promote<float2>(x)", "FETCH_SIZE", 2 * copy_bytes / 1024, 3, 10)
             + _counter_rows("forge::k_x(y)", "FETCH_SIZE", 2 * 5e9 / 1024, 2, 10, 10))
    write = (_counter_rows("forge::k_bt_promote<float2>(x)", "WRITE_SIZE", n * m * eb * K / 1024, 3, 10)
             + _counter_rows("forge::k_x(y)", "WRITE_SIZE", 1e9 / 1024, 2, 10, 10))
    _write_csv(tmp_path / "f.csv", fetch)
    _write_csv(tmp_path / "w.csv", write)
    lines = cs.traffic(str(tmp_path / "f.csv"), str(tmp_path / "w.csv"), n, K, eb)
    assert "reads = FETCH_SIZE x 0.5000" in lines[0] and "writes = WRITE_SIZE x 1.0000" in lines[0]
    row = next(r for r in lines if r.startswith("k_x"))
    calls, rd, wr = row.split()[1:]
    # two dispatches of 5 GB read and 1 GB written each
    assert calls == "2" and float(rd) == pytest.approx(10.0) and float(wr) == pytest.approx(2.0)


def test_busy_fraction(tmp_path):
    # 8 XCDs' GRBM_GUI_ACTIVE summed: busy = MFMA busy cycles x 8 / (GRBM x 1024 SIMDs)
    rows = (_counter_rows("forge::k_m(z)", "SQ_VALU_MFMA_BUSY_CYCLES", 512 * 1000, 1, 2000)
            + _counter_rows("forge::k_m(z)", "GRBM_GUI_ACTIVE", 8 * 1000, 1, 2000)
            + _counter_rows("forge::k_m(z)", "SQ_INSTS_LDS", 100, 1, 2000)
            + _counter_rows("forge::k_m(z)", "SQ_LDS_BANK_CONFLICT", 50, 1, 2000))
    _write_csv(tmp_path / "b.csv", rows)
    lines = cs.busy(str(tmp_path / "b.csv"))
    f = lines[1].split()
    assert f[0] == "k_m" and f[1] == "1" and float(f[3]) == pytest.approx(0.5) and float(f[4]) == pytest.approx(0.5)


def test_trace_summary_rates(tmp_path):
    n, K = 97, 1  # 3 groups of 32 sweeps
    m = (n + 2 * (min(20, n) // 2)) // 2
    G = (n - 1 + 31) // 32
    d = tmp_path / "trace"
    d.mkdir()
    _write_csv(d / "run_kernel_stats.csv", [
        {"Name": "void forge::k_ts_q2<double, forge::z2>(long)", "Calls": G, "TotalDurationNs": 3000,
         "AverageNs": 1000},
    ])
    # one back-transformation, launched group G - 1 first, 1 us each
    _write_csv(d / "run_kernel_trace.csv", [
        {"Kernel_Name": "void forge::k_ts_q2<double, forge::z2>(long)", "Dispatch_Id": q,
         "Start_Timestamp": 0, "End_Timestamp": 1000} for q in range(G)])
    bench = tmp_path / "b.json"
    bench.write_text(json.dumps({"value": 1.0, "steps": 1, "roofline": None,
                                 "config": {"iterations_by_phase": {"fp64": 1}, "converged": True}}))
    out = tmp_path / "summary.txt"
    argv = sys.argv
    sys.argv = ["trace_summary", str(d), str(bench), str(out), str(n), str(K)]
    try:
        ts.main()
    finally:
        sys.argv = argv
    text = out.read_text()
    # 16 flops per reflector element per eigenvector column over all n - 1 sweeps
    elems = sum(n - 1 - s for s in range(n - 1))
    rate = 16.0 * elems * m * K / (G * 1000e-9) / 1e12
    assert f"FP64 launches in the run: {G} (1 back-transformations of {G})" in text
    assert f"{rate:.2f} TFLOP/s" in text
    assert not d.exists()  # the raw trace is removed after the reduction
