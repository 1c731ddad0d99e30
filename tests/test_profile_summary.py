"""The rocprofv3 CSV summaries: duration statistics, per-dispatch trace
statistics, and calibrated HBM traffic, on small synthetic CSV files with
known answers."""
import csv
import json

import pytest

from tressfx_amd import profile_summary as ps


def _write(path, header, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(rows)
    return str(path)


def test_trace_stats_all_and_last_n(tmp_path):
    k16 = "void tfx::fast::k_fused_substep<16, 512>(tfx::DevHair const*)"
    other = "void tfx::exact::k_sdf_construct_multi<16>(tfx::SdfJob const*)"
    rows = []
    # dispatch ids out of order in the file; durations 1..6 us for k16
    for did, dur in ((5, 5), (1, 1), (3, 3), (6, 6), (2, 2), (4, 4)):
        rows.append([k16, did, 1000 * did, 1000 * did + 1000 * dur])
    rows.append([other, 7, 0, 9000])
    p = _write(tmp_path / "t.csv", ["Kernel_Name", "Dispatch_Id", "Start_Timestamp", "End_Timestamp"], rows)
    (rec,) = ps.trace_stats(p, "k_fused_substep<16", last_n=2)
    assert rec["name"] == k16
    assert rec["all"] == {"calls": 6, "avg_us": 3.5, "min_us": 1.0, "max_us": 6.0}
    assert rec["last_2"] == {"calls": 2, "avg_us": 5.5, "min_us": 5.0, "max_us": 6.0}
    assert "last_9" not in ps.trace_stats(p, "k_fused_substep<16", last_n=9)[0]


def test_kernel_stats(tmp_path):
    p = _write(tmp_path / "s.csv", ["Name", "Calls", "TotalDurationNs", "AverageNs", "MinNs", "MaxNs"],
               [["k_a", 4, 4000, 1000, 900, 1100], ["copy", 1, 10, 10, 10, 10]])
    (rec,) = ps.kernel_stats(p, "k_a")
    assert rec == {"name": "k_a", "calls": 4, "avg_us": 1.0, "min_us": 0.9, "max_us": 1.1, "total_ms": 0.004}


def test_traffic_calibrated_by_the_copy(tmp_path):
    copy_bytes = 1 << 20
    hdr = ["Kernel_Name", "Dispatch_Id", "Counter_Name", "Counter_Value"]
    # FETCH_SIZE sees half the copy's read bytes (factor 0.5); two counter
    # instances per dispatch are summed
    fetch = [["k_x", 1, "FETCH_SIZE", 100.0], ["k_x", 1, "FETCH_SIZE", 100.0], ["k_x", 2, "FETCH_SIZE", 300.0]]
    fetch += [["__amd_rocclr_copyBuffer", 10 + i, "FETCH_SIZE", 512.0] for i in range(3)]
    fetch += [["__amd_rocclr_copyBuffer", 20, "FETCH_SIZE", 1.0]]      # a small copy: not a calibration copy
    write = [["k_x", 1, "WRITE_SIZE", 50.0], ["k_x", 2, "WRITE_SIZE", 150.0]]
    write += [["__amd_rocclr_copyBuffer", 10 + i, "WRITE_SIZE", 1024.0] for i in range(3)]
    pf = _write(tmp_path / "f.csv", hdr, fetch)
    pw = _write(tmp_path / "w.csv", hdr, write)
    r = ps.traffic(pf, pw, "k_x", copy_bytes)
    assert r["FETCH_SIZE"]["factor"] == pytest.approx(0.5)
    assert r["FETCH_SIZE"]["kib_per_dispatch_mean"] == pytest.approx(250.0)
    assert r["FETCH_SIZE"]["bytes_per_dispatch"] == pytest.approx(250.0 * 1024 / 0.5)
    assert r["WRITE_SIZE"]["factor"] == pytest.approx(1.0)
    assert r["bytes_per_launch"] == pytest.approx(500.0 * 1024 + 100.0 * 1024)
    assert r["source_digest"] == ps.fused_source_digest()
    json.dumps(r)


def test_traffic_last_n_takes_the_last_dispatches(tmp_path):
    hdr = ["Kernel_Name", "Dispatch_Id", "Counter_Name", "Counter_Value"]
    cal = [["__amd_rocclr_copyBuffer", 100 + i, "FETCH_SIZE", 1024.0] for i in range(3)]
    rows = [["k_x", 9, "FETCH_SIZE", 30.0], ["k_x", 1, "FETCH_SIZE", 1000.0], ["k_x", 5, "FETCH_SIZE", 10.0]] + cal
    p = _write(tmp_path / "f.csv", hdr, rows)
    pw =_write(tmp_path / "w.csv", hdr, [[n, d, "WRITE_SIZE", v] for n, d, _, v in rows])
    r = ps.traffic(p, pw, "k_x", 1 << 20, last_n=2)
    assert r["FETCH_SIZE"]["kib_per_dispatch_mean"] == pytest.approx(20.0)     # dispatches 5 and 9
    with pytest.raises(SystemExit):
        ps.traffic(p, pw, "k_x", 1 << 20, last_n=4)


def test_traffic_needs_a_calibration_copy(tmp_path):
    hdr = ["Kernel_Name", "Dispatch_Id", "Counter_Name", "Counter_Value"]
    pf = _write(tmp_path / "f.csv", hdr, [["k_x", 1, "FETCH_SIZE", 1.0]])
    with pytest.raises(SystemExit):
        ps.traffic(pf, pf, "k_x", 1 << 20)


def test_counter_ratios(tmp_path):
    hdr = ["Kernel_Name", "Dispatch_Id", "Counter_Name", "Counter_Value"]
    rows = [["k_x", 1, "SQ_WAVES", 10.0], ["k_x", 1, "SQ_INSTS_VALU", 1000.0], ["k_x", 1, "SQ_WAVE_CYCLES", 400.0],
            ["k_x", 1, "SQ_ACTIVE_INST_VALU", 100.0], ["k_x", 1, "SQ_ACTIVE_INST_LDS", 8.0],
            ["k_x", 1, "SQ_LDS_BANK_CONFLICT", 2.0]]
    p = _write(tmp_path / "c.csv", hdr, rows)
    r = ps.counters(p, "k_x")["ratios"]
    assert r["valu_insts_per_wave"] == pytest.approx(100.0)
    assert r["valu_issue_per_wave_cycle"] == pytest.approx(0.25)
    assert r["lds_issue_per_wave_cycle"] == pytest.approx(0.02)
    assert r["lds_bank_conflict_per_lds_issue_cycle"] == pytest.approx(0.25)
