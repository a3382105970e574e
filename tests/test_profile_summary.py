"""CPU tests of the rocprofv3 summariser behind profiles/ (and bench.py's
roofline traffic): synthetic CSVs in rocprofv3's column layout, each figure
checked against a hand computation."""
import csv
import re

import pytest

from distembed_amd import profile_summary as ps

COUNTER_FIELDS = ["Correlation_Id", "Dispatch_Id", "Agent_Id", "Queue_Id", "Process_Id",
                  "Thread_Id", "Grid_Size", "Kernel_Id", "Kernel_Name", "Workgroup_Size",
                  "LDS_Block_Size", "Scratch_Size", "VGPR_Count", "Accum_VGPR_Count",
                  "SGPR_Count", "Counter_Name", "Counter_Value", "Start_Timestamp",
                  "End_Timestamp"]


def _counters(path, rows):
    """rows: (dispatch, kernel, counter, value, start_ns, end_ns)."""
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=COUNTER_FIELDS)
        w.writeheader()
        for d, k, c, v, t0, t1 in rows:
            w.writerow({"Correlation_Id": d, "Dispatch_Id": d, "Kernel_Name": k,
                        "Counter_Name": c, "Counter_Value": v, "Start_Timestamp": t0,
                        "End_Timestamp": t1})


def _value(text, key):
    m = re.search(re.escape(key) + r"\s*=\s*([0-9.e+-]+)", text)
    assert m, key
    return float(m.group(1))


def test_trace_averages_force_launches_only(tmp_path):
    stats = tmp_path / "stats.csv"
    with open(stats, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Name", "Calls", "TotalDurationNs", "AverageNs"])
        w.writeheader()
        w.writerow({"Name": "fe::force_kernel1<1, 0, 16>(fe::StepDesc)", "Calls": 2,
                    "TotalDurationNs": 30000, "AverageNs": 15000})
    trace = tmp_path / "trace.csv"
    with open(trace, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        w.writeheader()
        for name, t0, t1 in (("fe::force_kernel1<1, 0, 16>(fe::StepDesc)", 0, 10000),
                             ("fe::force_kernel1<2, 1, 16>(fe::StepDesc)", 0, 20000),
                             ("fe::epoch_kernel(int)", 0, 99000)):
            w.writerow({"Kernel_Name": name, "Start_Timestamp": t0, "End_Timestamp": t1})
    out = tmp_path / "out.txt"
    ps.trace(str(stats), str(trace), str(out), "cmd")
    text = out.read_text()
    # 10 us and 20 us force launches; the epoch kernel is not a force launch
    assert "force launches, all instantiations: n=2 avg_us=15.00" in text
    ps.trace(str(stats), str(trace), str(out), "cmd", r"force_kernel1?<\d+, 1,")
    assert "force launches, all instantiations: n=1 avg_us=20.00" in out.read_text()


def test_pmc_scales_by_the_copy_calibration(tmp_path):
    gib_kib = (1 << 30) / 1024
    calf, calw = tmp_path / "calf.csv", tmp_path / "calw.csv"
    # the copy kernel's counters read half (fetch) and all (write) of the copied KiB
    _counters(calf, [(1, "__amd_rocclr_copyBuffer", "FETCH_SIZE", gib_kib / 2, 0, 1)])
    _counters(calw, [(1, "__amd_rocclr_copyBuffer", "WRITE_SIZE", gib_kib, 0, 1)])
    fetch, write = tmp_path / "f.csv", tmp_path / "w.csv"
    k1, k2 = "fe::force_kernel1<1, 0, 16>", "fe::force_kernel1<2, 0, 16>"
    _counters(fetch, [(1, k1, "FETCH_SIZE", 100.0, 0, 1), (2, k1, "FETCH_SIZE", 300.0, 0, 1),
                      (3, k2, "FETCH_SIZE", 50.0, 0, 1),
                      (4, "fe::epoch_kernel", "FETCH_SIZE", 1e6, 0, 1)])
    _counters(write, [(1, k1, "WRITE_SIZE", 10.0, 0, 1), (2, k1, "WRITE_SIZE", 10.0, 0, 1),
                      (3, k2, "WRITE_SIZE", 40.0, 0, 1)])
    out = tmp_path / "out.txt"
    ps.pmc(str(fetch), str(write), str(out), "cmd", str(calf), str(calw))
    text = out.read_text()
    # reads: KiB x 1024 / 0.5 per dispatch, averaged over the three force dispatches
    read = (100.0 + 300.0 + 50.0) * 1024 / 0.5 / 3
    wr = (10.0 + 10.0 + 40.0) * 1024 / 1.0 / 3
    assert _value(text, "hbm_read_bytes_per_launch") == pytest.approx(read, rel=1e-6)
    assert _value(text, "hbm_write_bytes_per_launch") == pytest.approx(wr, rel=1e-6)
    assert _value(text, "hbm_bytes_per_launch") == pytest.approx(read + wr, rel=1e-6)


def test_issue_valu_busy_per_batch(tmp_path):
    sq = tmp_path / "sq.csv"
    k16, k32 = "fe::force_kernel1<2, 1, 16>", "fe::force_kernel1<2, 0, 16>"
    rows = []
    d = 0
    # two epochs of two batches of the BF16 form, plus an FP32 launch to ignore
    for ep in range(2):
        for b, (waves, valu, us) in enumerate(((1000, 600_000, 10.0), (500, 60_000, 5.0))):
            d += 1
            t1 = int(us * 1000)
            rows += [(d, k16, "SQ_WAVES", waves, 0, t1), (d, k16, "SQ_INSTS_VALU", valu, 0, t1),
                     (d, k16, "SQ_INSTS_VMEM_RD", 4 * waves, 0, t1)]
    d += 1
    rows.append((d, k32, "SQ_WAVES", 7, 0, 1))
    _counters(sq, rows)
    out = tmp_path / "out.txt"
    ps.issue(str(sq), str(out), "cmd", 2, 1)
    text = out.read_text()
    assert "BF16 gathers: 2 epoch(s) of 2 launches" in text
    lines = [l.split() for l in text.splitlines() if re.match(r"\s+\d+\s+\d+", l)]
    assert [int(l[0]) for l in lines] == [0, 1]
    # batch 0: 600k VALU x 4 cycles / 1024 SIMDs / (10 us x 2400 cycles/us)
    busy0 = 600_000 * 4 / 1024 / (10.0 * 2400)
    busy1 = 60_000 * 4 / 1024 / (5.0 * 2400)
    assert float(lines[0][2]) == pytest.approx(600.0)   # VALU per wave
    assert float(lines[0][3]) == pytest.approx(4.0)     # vector reads per wave
    assert float(lines[0][5]) == pytest.approx(busy0, abs=0.005)
    assert float(lines[1][5]) == pytest.approx(busy1, abs=0.005)
    # the FP32 form holds one dispatch, less than an epoch of two batches
    with pytest.raises(SystemExit):
        ps.issue(str(sq), str(out), "cmd", 2, 0)
