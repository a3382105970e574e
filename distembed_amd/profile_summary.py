r"""Summarise rocprofv3 CSV output into the short text kept under profiles/.

  python -m distembed_amd.profile_summary trace  <kernel_stats.csv> <kernel_trace.csv> <out> <command> [<regex>]
  python -m distembed_amd.profile_summary pmc <fetch csv> <write csv> <out> <command> <copyprobe fetch csv> <copyprobe write csv> [<regex>]
  python -m distembed_amd.profile_summary issue <counter csv> <out> <command> <batches> <form 0|1> [<tcc csv>]
  python -m distembed_amd.profile_summary copyprobe

trace: per-kernel launch count, total and average duration (from the
--kernel-trace --stats run), plus the force kernel's per-launch average
(only the instantiations whose name matches <regex>, when given: e.g.
"force_kernel1?<\d+, 1," for the BF16-gather ones).
pmc: FETCH_SIZE and WRITE_SIZE (KiB per dispatch) of the force kernel (or
of the kernels <regex> selects, e.g. "spmm_kernel<\d+, 0,"), scaled by what
the same counters read for a cold 1 GiB device-to-device copy (copyprobe),
and the resulting HBM bytes per launch (hbm_bytes_per_launch).
issue: per batch index, for one gather form (force_kernel1<U, 0|1, G>:
0 FP32, 1 BF16) of a run with one tiling of it, the SQ counters of an
--pmc run of whole epochs (waves, VALU and vector-memory-read
instructions per wave) and the VALU busy fraction: VALU instructions x 4
cycles (a wave64 VALU instruction occupies a 16-lane SIMD for 4 cycles),
over 1024 SIMDs (256 CUs x 4) and the dispatch's duration at 2.4 GHz (the
peak engine clock, so the fraction is a lower bound); with a second csv of
TCC_HIT_sum / TCC_MISS_sum, the L2 hit rate.
"""
from __future__ import annotations

import csv
import re
import statistics
import sys
from collections import defaultdict


def _rows(path: str) -> list:
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def _write(path: str, lines: list) -> None:
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def _short(name: str) -> str:
    return name.split("(")[0].replace("void ", "").strip()


def trace(stats_csv: str, trace_csv: str, out: str, command: str, match: str = "") -> None:
    rows = list(_rows(stats_csv))
    lines = [f"# rocprofv3 kernel statistics", f"command: {command}", ""]
    lines.append(f"{'kernel':60s} {'calls':>7s} {'total_ms':>10s} {'avg_us':>9s}")
    for r in rows:
        name = _short(r.get("Name", r.get("KernelName", "?")))
        calls = int(float(r.get("Calls", 0)))
        tot = float(r.get("TotalDurationNs", 0)) / 1e6
        avg = float(r.get("AverageNs", 0)) / 1e3
        lines.append(f"{name[:60]:60s} {calls:7d} {tot:10.3f} {avg:9.2f}")
    durs = defaultdict(list)
    for r in _rows(trace_csv):
        durs[_short(r["Kernel_Name"])].append(
            (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    lines.append("")
    allv = []
    for k, v in sorted(durs.items()):
        if "force_kernel" in k and (not match or re.search(match, k)):
            allv += v
            lines.append(f"force launches {k}: n={len(v)} avg_us={statistics.mean(v):.2f} "
                         f"median_us={statistics.median(v):.2f} max_us={max(v):.2f}")
    if allv:
        lines.append(f"force launches, all instantiations: n={len(allv)} "
                     f"avg_us={statistics.mean(allv):.2f} median_us={statistics.median(allv):.2f} "
                     f"max_us={max(allv):.2f}")
    _write(out, lines)


def _per_dispatch(path: str, counter: str):
    vals = defaultdict(dict)  # kernel -> dispatch -> value
    for r in _rows(path):
        if r.get("Counter_Name") != counter:
            continue
        vals[_short(r["Kernel_Name"])][int(r["Dispatch_Id"])] = float(r["Counter_Value"])
    return vals


def pmc(fetch_csv: str, write_csv: str, out: str, command: str, cal_fetch_csv: str,
        cal_write_csv: str, match: str = "") -> None:
    fetch = _per_dispatch(fetch_csv, "FETCH_SIZE")
    write = _per_dispatch(write_csv, "WRITE_SIZE")
    cal_f = calibration(cal_fetch_csv, "FETCH_SIZE")
    cal_w = calibration(cal_write_csv, "WRITE_SIZE")
    lines = ["# rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (KiB per dispatch)",
             f"command (fetch): rocprofv3 --pmc FETCH_SIZE --output-format csv -- {command}",
             f"command (write): rocprofv3 --pmc WRITE_SIZE --output-format csv -- {command}",
             "calibration: rocprofv3 --pmc <counter> -- python3 -m distembed_amd.profile_summary "
             "copyprobe (a cold 1 GiB device-to-device copy)", ""]
    lines.append(f"calibration factors: FETCH_SIZE reads {cal_f:.4f} x the copied KiB, "
                 f"WRITE_SIZE reads {cal_w:.4f} x")
    tot_f = tot_w = 0.0
    tot_n = 0
    for k in sorted(fetch):
        # the kernels the regex selects (default: every force-kernel instantiation)
        if not (re.search(match, k) if match else "force_kernel" in k):
            continue
        f = statistics.mean(fetch[k].values())
        w = statistics.mean(write.get(k, {0: 0.0}).values())
        fb = f * 1024 / (cal_f or 1.0)
        wb = w * 1024 / (cal_w or 1.0)
        n = len(fetch[k])
        tot_f += fb * n
        tot_w += wb * n
        tot_n += n
        lines.append(f"{k}: dispatches={n} fetch_KiB={f:.1f} write_KiB={w:.1f} "
                     f"read_bytes={fb:.4e} write_bytes={wb:.4e}")
    if tot_n:
        # every selected instantiation (the force step's batches use several)
        lines.append(f"all selected launches: dispatches={tot_n}")
        lines.append(f"hbm_bytes_per_launch = {(tot_f + tot_w) / tot_n:.6e}")
        lines.append(f"hbm_read_bytes_per_launch = {tot_f / tot_n:.6e}")
        lines.append(f"hbm_write_bytes_per_launch = {tot_w / tot_n:.6e}")
    _write(out, lines)


def copyprobe(nbytes: int = 1 << 30) -> None:
    """A plain device-to-device copy of `nbytes` cold bytes (the HIP runtime's
    copy kernel), run under rocprofv3 --pmc to calibrate FETCH_SIZE and
    WRITE_SIZE: the counters should read nbytes / 1024 KiB each."""
    import ctypes as C
    from . import _lib
    _lib.load()
    hip = C.CDLL("libamdhip64.so.7")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    # flush the caches between the fill and the copy with an unrelated copy
    c, d2 = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(c), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(d2), C.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(d2, c, C.c_size_t(nbytes), 3) == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    print(f"copyprobe: copied {nbytes} bytes device to device", flush=True)


def issue(csv_path: str, out: str, command: str, batches: int, form: int,
          tcc_csv: str = "") -> None:
    def per_form(path):
        cnt = defaultdict(dict)
        info = {}
        for r in _rows(path):
            m = re.search(r"force_kernel1<\d+, (\d+), \d+>", r["Kernel_Name"])
            if not m:
                continue
            d = int(r["Dispatch_Id"])
            cnt[d][r["Counter_Name"]] = float(r["Counter_Value"])
            info[d] = (int(m.group(1)),
                       (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3)
        forms = defaultdict(list)
        for d in sorted(cnt):
            forms[info[d][0]].append((cnt[d], info[d][1]))
        return forms
    sq = per_form(csv_path)
    tcc = per_form(tcc_csv) if tcc_csv else {}
    simds, clock_ghz = 1024, 2.4
    lines = ["# rocprofv3 --pmc SQ counters of the force kernel, per batch index",
             f"command: rocprofv3 --pmc SQ_WAVES SQ_INSTS_VALU SQ_INSTS_VMEM_RD SQ_ACTIVE_INST_VALU "
             f"SQ_WAIT_INST_ANY SQ_BUSY_CYCLES SQ_WAVE_CYCLES SQ_ACTIVE_INST_ANY -- {command}"]
    if tcc_csv:
        lines.append(f"command (L2): rocprofv3 --pmc TCC_HIT_sum TCC_MISS_sum GRBM_GUI_ACTIVE -- {command}")
    lines += ["valu_busy = SQ_INSTS_VALU x 4 cycles / 1024 SIMDs / (dispatch us x 2400 cycles/us); "
              "durations are those of the counter run", ""]
    for name in ("FP32 gathers" if form == 0 else "BF16 gathers",):
        rows = sq.get(form, [])
        if not rows:
            raise SystemExit(f"no force_kernel1 dispatches of form {form} in {csv_path}")
        n_ep = len(rows) // batches
        if n_ep < 1:
            raise SystemExit(f"{len(rows)} dispatches of form {form} in {csv_path}: "
                             f"fewer than one epoch of {batches}")
        lines.append(f"{name}: {n_ep} epoch(s) of {batches} launches")
        lines.append("batch  waves   valu/wave  vmem_rd/wave  us      valu_busy" +
                     ("  l2_hit" if tcc else ""))
        tot_v = tot_t = 0.0
        for b in range(batches):
            sel = [rows[e * batches + b] for e in range(n_ep)]
            w = statistics.mean(c["SQ_WAVES"] for c, _ in sel)
            v = statistics.mean(c["SQ_INSTS_VALU"] for c, _ in sel)
            vm = statistics.mean(c["SQ_INSTS_VMEM_RD"] for c, _ in sel)
            us = statistics.median(t for _, t in sel)
            busy = v * 4 / simds / (us * clock_ghz * 1e3)
            tot_v += v * 4 / simds / (clock_ghz * 1e3)
            tot_t += us
            line = f"{b:5d} {w:7.0f} {v / w:10.1f} {vm / w:12.2f} {us:8.1f} {busy:9.2f}"
            if tcc:
                tsel = [tcc[form][e * batches + b][0] for e in range(len(tcc[form]) // batches)]
                h = sum(c["TCC_HIT_sum"] for c in tsel)
                mi = sum(c["TCC_MISS_sum"] for c in tsel)
                line += f"  {h / max(h + mi, 1.0):6.2f}"
            lines.append(line)
        lines.append(f"epoch: {tot_t:.1f} us of dispatches, VALU {tot_v:.1f} us at the peak clock "
                     f"({tot_v / tot_t:.2f})")
        lines.append("")
    _write(out, lines)


def calibration(path: str, counter: str, nbytes: int = 1 << 30) -> float:
    """Counter KiB of the last runtime copy kernel / the KiB it copied."""
    vals = []
    for r in _rows(path):
        if r.get("Counter_Name") == counter and "copyBuffer" in r["Kernel_Name"]:
            vals.append((int(r["Dispatch_Id"]), float(r["Counter_Value"])))
    vals.sort()
    return vals[-1][1] / (nbytes / 1024.0)


if __name__ == "__main__":
    mode = sys.argv[1]
    if mode == "trace":
        trace(sys.argv[2], sys.argv[3], sys.argv[4], sys.argv[5],
              sys.argv[6] if len(sys.argv) > 6 else "")
    elif mode == "copyprobe":
        copyprobe()
    elif mode == "issue":
        issue(sys.argv[2], sys.argv[3], sys.argv[4], int(sys.argv[5]), int(sys.argv[6]),
              sys.argv[7] if len(sys.argv) > 7 else "")
    else:
        pmc(sys.argv[2], sys.argv[3], sys.argv[4], sys.argv[5], sys.argv[6], sys.argv[7],
            sys.argv[8] if len(sys.argv) > 8 else "")
