#!/usr/bin/env python3
"""Throughput of the force-model embedding hot path on one MI355X.

Workload: R-MAT scale 20 (edge factor 16, a=0.57, b=c=0.19; symmetrised,
duplicates and self-loops removed, natural vertex order), d=128, batch 16384,
5 reference negatives per batch, lr 0.02 -- the one-GPU configuration the
metric is quoted on. A step is one epoch: the 64 dependent minibatches of
algo_force2_vec_ns over the whole graph, replayed as captured HIP graphs.
Inputs are resident in HBM before the timed region starts. The launch
tiling is the per-batch tiling the library's autotuner (fe_autotune) chose
for this workload, kept under profiles/ with the rocprofv3 summaries taken
with it (config.tuning_source); --autotune searches afresh instead.

value = (processed CSR entries + (row, negative) pairs) / second, FP32.
roofline: the force kernel's algorithmic bytes per launch (its share of the
per-epoch byte count: gathered rows d*4+4 per processed entry, own row read
and new row written d*4 each, row pointers, the batch's negative rows)
divided by its average launch duration: one HIP event pair on its stream
around the timed epochs, divided by the force launches in them (one 4 us
epoch tick per 64 launches included). A second pass with an event pair
around every launch is reported beside it; those event nodes add about
2 us per launch. traffic: HBM bytes per launch from the
committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary, when present.
bf16_form: the same workload with BF16 gathers (FP32 master and
accumulation), its speed-up and its final loss against the FP32 run, and
against an FP32 run with the same tiling (the tiling alone moves the final
loss of these chaotic runs; that spread is reported too); loss_over_seeds:
both forms from 8 further initial embeddings, the mean per-seed relative
difference of the final losses and its standard error.
cpu_baseline: the oracle's FP64 restatement of the reference loop (OpenMP,
schedule(static)) on a bounded sample of the same workload.
precision: R-MAT scale 20 and 22 at 5 and 10 negatives, FP32 and BF16
(each autotuned): epoch times, BF16 speed-ups, final losses (also like for
like, at the BF16 run's tiling) and per-step gradient errors of the BF16
form against the FP32 kernel; for R-MAT-22 at 5 negatives, the final
losses from 4 initial embeddings as in loss_over_seeds (about 10 s each).
serving: Poisson arrivals of R-MAT embedding jobs (2e4-2e7 edges, d from
{32, 64, 128}, 50 epochs) at 4, 8 and 16 jobs/s under an 8 GiB budget:
p50/p99 job latency (arrival to result on the host) and edges+negatives/s.
batching: the 512-graph SBM pool (d=64, B=1024, 50 epochs) in captured
rounds, FP32 and BF16, tiles from the library's engine autotuner.

--form precision | batching | serving runs the other workloads instead
(distembed_amd/bench_forms.py) and prints their results as one JSON line;
--form spmm times the SpMM sibling (fe_spmm) on the same R-MAT graph.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "edges+negatives/sec at d=128 vs plain GPU; embed-job p50/p99 latency at QPS"
HBM_PEAK_GBPS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def traffic_from_profile(name: str):
    """HBM bytes per force launch from the committed rocprofv3 --pmc summary."""
    p = ROOT / "profiles" / name
    if not p.exists():
        return None
    m = re.search(r"hbm_bytes_per_launch\s*[:=]\s*([0-9.eE+]+)", p.read_text())
    return float(m.group(1)) if m else None


def load_tuning(arg: str):
    """--tuning value: JSON text, @path of a JSON file, or empty (None)."""
    if not arg:
        return None
    if arg.startswith("@"):
        return json.loads(Path(arg[1:]).read_text())
    return json.loads(arg)


def cpu_baseline(g, lists, d, B, ns, lr, budget_s: float, max_epochs: int):
    sys.path.insert(0, str(ROOT / "tests"))
    import _oracle as orc  # the oracle, timed as the CPU baseline
    from distembed_amd import api
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    X = api.init_uniform(g.n, d, -0.5, 0.5, seed=0).astype(np.float64)
    batches = -(-g.n // B)
    units_epoch = int((g.col_idx < g.n - 1).sum()) + batches * B * ns
    done, t_total = 0, 0.0
    while done < max_epochs and t_total < budget_s:
        t0 = time.perf_counter()
        orc.train(g, X, B, 1, lists[done:], lr=lr, quirks=3, threads=threads)
        t_total += time.perf_counter() - t0
        done += 1
        log(f"cpu baseline: epoch {done} done, {t_total:.1f} s")
    return {"value": units_epoch * done / t_total, "unit": "edges+negatives/s",
            "cores": threads, "kind": "port",
            "sample": f"{done} epoch(s) of the same R-MAT-20 workload, FP64 (the reference's "
                      f"VALUE_TYPE), {threads} OpenMP threads, {t_total:.1f} s"}


def spmm_form(api, args) -> dict:
    """The SpMM sibling (fe_spmm, SpMMAlgo::algo_spmm): Y = lr * A' X on the
    R-MAT graph at d = args.d, FP32 and BF16 gathers, B = args.batch. One
    repetition is one full product (every batch); its time is the difference
    between (steps + 1) repetitions and one (the shortest of three calls
    each), which removes the final download. Algorithmic bytes: processed entries x (row + column id),
    row pointers, and the output rows written. The CPU baseline is the
    oracle's FP64 product on the host cores (one repetition)."""
    t0 = time.perf_counter()
    g = api.Graph.rmat(args.scale, 16, 0.57, 0.19, 0.19, seed=0)
    log(f"spmm: graph n={g.n} nnz={g.nnz} ({time.perf_counter() - t0:.1f} s)")
    d, B, lr, k = args.d, args.batch, args.lr, max(args.steps, 1)
    X = api.init_uniform(g.n, d, -0.5, 0.5, seed=0)
    edges = int((g.col_idx < g.n - 1).sum())
    res = {"metric": "SpMM edges/s (lr * A' X, d=%d)" % d, "form": "spmm", "unit": "edges/s",
           "higher_is_better": True, "n_gpus": 1, "data": "synthetic", "steps": k,
           "config": {"workload": f"rmat{args.scale}_d{d}_spmm", "n": g.n, "nnz": g.nnz,
                      "edges_processed": edges, "d": d, "batch": B, "lr": lr}}
    for name, gather, s in (("fp32", api.FP32, 4), ("bf16", api.BF16, 2)):
        alg = edges * (d * s + 4) + (g.n + 1) * 8 + g.n * d * 4
        with api.ForceEmbed(g.n, d, gather=gather) as ctx:
            ctx.upload_csr(g)
            ctx.upload_x(X)
            ctx.configure(B, lr, 5.0)
            ctx.spmm(2, B, lr)  # plan and warm-up
            one, many = [], []
            for _ in range(3):  # the shortest of three of each: the downloads vary
                t1 = time.perf_counter()
                ctx.spmm(1, B, lr)
                t2 = time.perf_counter()
                ctx.spmm(k + 1, B, lr)
                t3 = time.perf_counter()
                one.append(t2 - t1)
                many.append(t3 - t2)
        per = (min(many) - min(one)) / k
        res[name] = {"value": edges / per, "ms_per_product": per * 1e3,
                     "roofline": {"bound": "hbm", "achieved": alg / per / 1e9, "peak": HBM_PEAK_GBPS,
                                  "unit": "GB/s", "frac": alg / per / 1e9 / HBM_PEAK_GBPS,
                                  "algorithmic_bytes": alg}}
        log(f"spmm {name}: {per * 1e3:.3f} ms per product, {edges / per / 1e9:.2f} G edges/s, "
            f"{alg / per / 1e9:.0f} GB/s algorithmic")
    res["value"] = res["fp32"]["value"]
    res["bf16_speedup"] = res["fp32"]["ms_per_product"] / res["bf16"]["ms_per_product"]
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "tests"))
        import _oracle as orc  # the oracle, timed as the CPU baseline
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
        X64 = X.astype(np.float64)
        t1 = time.perf_counter()
        orc.spmm(g, X64, lr, quirks=3, threads=threads)
        dt = time.perf_counter() - t1
        res["cpu_baseline"] = {"value": edges / dt, "unit": "edges/s", "cores": threads,
                               "kind": "port",
                               "sample": f"one product on the same graph, FP64, {threads} "
                                         f"OpenMP threads, {dt:.1f} s"}
    return res


def run_form(api, g, X0, lists, gather, args, tuning=None):
    """Autotune, warm up, time args.steps epochs; returns the measurements."""
    d, B, ns, lr = args.d, args.batch, args.ns, args.lr
    ctx = api.ForceEmbed(g.n, d, gather=gather)
    try:
        ctx.upload_csr(g)
        ctx.upload_x(X0)
        ctx.set_negatives(lists)
        ctx.configure(B, lr, 5.0)
        if tuning:
            ctx.set_tuning(**tuning.get("tuning", {}))
            if tuning.get("per_batch"):
                ctx.set_batch_tunings(tuning["per_batch"])
        elif not args.no_autotune:
            best, best_ms, table = ctx.autotune(epochs=2, search_chunking=True)
            log(f"autotune ({'bf16' if gather else 'fp32'}): {best} {best_ms:.3f} ms/epoch "
                f"over {len(table)} trials")
        tun = ctx.get_tuning()
        per_batch = [[t[k] for k in ("unroll", "waves_per_block", "items_per_wave", "chunk",
                                     "max_chunks", "light", "slice")] for t in ctx.batch_tunings()]
        edges, n0, n1 = ctx.epoch_units(ns)
        use_graph = not args.no_graph
        if args.warmup:
            ctx.train(args.warmup, B, ns, lr, use_graph=use_graph)
        # throughput pass: plain graph replays, every form from the same X0,
        # with one HIP event pair on the context's stream around the timed
        # epochs (the force launches' average over the timed region)
        ctx.upload_x(X0)
        ctx.set_timing(2)
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.train(args.steps, B, ns, lr, use_graph=use_graph)
        ctx.synchronize()
        dt = time.perf_counter() - t0
        kt_region = ctx.kernel_times()
        ctx.set_timing(False)
        attr, rep = ctx.loss(0, ns)
        # kernel pass: the same epochs again from X0 with a HIP event pair on
        # the context's stream around every force launch (reading the events
        # costs a sync per epoch, so this pass is not the throughput figure)
        ctx.upload_x(X0)
        ctx.set_timing(True)
        ctx.synchronize()
        t1 = time.perf_counter()
        ctx.train(args.steps, B, ns, lr, use_graph=use_graph)
        ctx.synchronize()
        dt_events = time.perf_counter() - t1
        kt = ctx.kernel_times()
        ctx.set_timing(False)
    finally:
        ctx.close()
    units = args.steps * edges + (n0 + (args.steps - 1) * n1 if args.steps else 0)
    return {"dt": dt, "dt_events": dt_events, "units": units, "edges": edges, "kt": kt,
            "kt_region": kt_region,
            "tuning": {"tuning": tun, "per_batch": per_batch}, "loss": attr + rep,
            "loss_parts": (attr, rep), "use_graph": use_graph}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed epochs")
    ap.add_argument("--warmup", type=int, default=3, help="untimed epochs")
    ap.add_argument("--scale", type=int, default=20)
    ap.add_argument("--d", type=int, default=128)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--ns", type=int, default=5)
    ap.add_argument("--lr", type=float, default=0.02)
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU baseline")
    ap.add_argument("--cpu-epochs", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-bf16", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--autotune", action="store_true",
                    help="search the tiling with fe_autotune even where a kept tiling exists")
    ap.add_argument("--tuning", type=str, default="",
                    help='json {"tuning": {fe_tuning fields}, "per_batch": [[unroll, '
                         'waves_per_block, items_per_wave, chunk, max_chunks, light] per '
                         'batch]} (as reported in config.tuning), or @file holding it; '
                         'skips the autotuner')
    ap.add_argument("--no-same-tiling", action="store_true",
                    help="skip the FP32 run with the BF16 run's tiling (the like-for-like loss)")
    ap.add_argument("--loss-seeds", type=int, default=8,
                    help="initial embeddings over which the BF16 and FP32 final losses are "
                         "compared (0: skip)")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-batching", action="store_true")
    ap.add_argument("--bf16-tuning", type=str, default="",
                    help="as --tuning, for the BF16 pass (skips its autotuner)")
    ap.add_argument("--form", choices=["plain", "precision", "batching", "serving", "spmm"],
                    default="plain")
    ap.add_argument("--epochs", type=int, default=50, help="epochs of the batching/serving jobs")
    ap.add_argument("--graphs", type=int, default=512, help="graphs of the batching form")
    ap.add_argument("--jobs", type=int, default=48, help="jobs per load level (serving form)")
    ap.add_argument("--qps", type=str, default="4,8,16", help="offered loads (serving form)")
    ap.add_argument("--budget-gib", type=float, default=0.0, help="engine HBM budget")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the library has no multi-GPU path "
              "(the halo exchange across GPUs is a design item); use --gpus 1", file=sys.stderr)
        return 2

    from distembed_amd import _lib, api
    if _lib.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        return 3

    if args.form == "spmm":
        print(json.dumps(spmm_form(api, args)), flush=True)
        return 0
    if args.form != "plain":
        from distembed_amd import bench_forms as bf
        if args.form == "precision":
            res = bf.precision(args.steps, args.warmup)
        elif args.form == "batching":
            res = bf.batching(args.epochs, count=args.graphs, budget_gib=args.budget_gib or 96.0)
        else:
            res = bf.serving(tuple(float(q) for q in args.qps.split(",")), args.jobs,
                             args.epochs, budget_gib=args.budget_gib or 8.0)
        res.update({"n_gpus": 1, "data": "synthetic", "unit": "edges+negatives/s"})
        print(json.dumps(res), flush=True)
        return 0

    t0 = time.perf_counter()
    g = api.Graph.rmat(args.scale, 16, 0.57, 0.19, 0.19, seed=0)
    log(f"graph: n={g.n} nnz={g.nnz} ({time.perf_counter() - t0:.1f} s)")
    d, B, ns, lr = args.d, args.batch, args.ns, args.lr
    batches = -(-g.n // B)
    X0 = api.init_uniform(g.n, d, -0.5, 0.5, seed=0)
    epochs_needed = max(args.steps, args.warmup, args.cpu_epochs, 3)
    lists, hits = api.reference_negatives(g.n, epochs_needed + batches - 1, ns)
    # the autotuner's tiling for this workload, as profiled (profiles/)
    kept = (args.scale, d, B, ns) == (20, 128, 16384, 5) and not args.autotune
    kept32 = ROOT / "profiles" / "force_kernel_rmat20_d128_fp32_tuning.json"
    kept16 = ROOT / "profiles" / "force_kernel_rmat20_d128_bf16_tuning.json"
    if not args.tuning and kept and kept32.exists():
        args.tuning = "@" + str(kept32)
    if not args.bf16_tuning and kept and kept16.exists():
        args.bf16_tuning = "@" + str(kept16)
    tuning = load_tuning(args.tuning)
    def source(a):
        return (str(Path(a[1:]).relative_to(ROOT)) if a.startswith("@") and
                Path(a[1:]).resolve().is_relative_to(ROOT) else ("--tuning" if a else "fe_autotune"))

    f32 = run_form(api, g, X0, lists, api.FP32, args, tuning)
    kt, kr = f32["kt"], f32["kt_region"]
    log(f"fp32: {args.steps} epochs in {f32['dt']:.3f} s; force {kr['force_ms'] * 1e3:.1f} us "
        f"per launch over the timed region ({kr['launches']} launches), "
        f"{kt['force_ms'] * 1e3:.1f} us with an event pair per launch; loss {f32['loss']:.6e}")
    value = f32["units"] / f32["dt"]

    edges = f32["edges"]
    per_epoch_bytes = (edges * (d * 4 + 4) + (g.n + 1) * 8 + 2 * g.n * d * 4
                       + batches * ns * d * 4)
    bytes_per_launch = per_epoch_bytes / batches
    force_s = kr["force_ms"] * 1e-3
    achieved = bytes_per_launch / force_s / 1e9 if force_s > 0 else None
    traffic = traffic_from_profile("force_kernel_rmat20_d128_fp32_pmc.txt")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "edges+negatives/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": f32["dt"] / max(args.steps, 1) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": f"rmat{args.scale}_d{d}_fp32",
            "graph": f"R-MAT scale {args.scale}, edge factor 16, a=0.57 b=c=0.19, seed 0, "
                     "symmetrised, deduplicated, no self-loops",
            "n": g.n, "nnz": g.nnz, "d": d, "batch": B, "ns": ns, "lr": lr,
            "epochs_timed": args.steps, "step": f"one epoch ({batches} dependent minibatches)",
            "graph_replay": f32["use_graph"], "negative_draws_equal_to_n": hits,
            "tuning": f32["tuning"],
            "tuning_source": source(args.tuning),
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBPS) if achieved else None,
            "traffic": traffic,
            "kernel": "force_kernel1",
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "avg_launch_ms": kr["force_ms"],
            "launches_timed": kr["launches"],
            "avg_launch_ms_event_pair_per_launch": kt["force_ms"],
            "event_pass_ms_per_step": f32["dt_events"] / max(args.steps, 1) * 1e3,
        },
        "loss_fp64_after_timed_epochs": f32["loss"],
    }
    if not args.no_bf16:
        b16 = run_form(api, g, X0, lists, api.BF16, args, load_tuning(args.bf16_tuning))
        kb = b16["kt_region"]
        log(f"bf16: {args.steps} epochs in {b16['dt']:.3f} s; force {kb['force_ms'] * 1e3:.1f} us "
            f"per launch over the timed region, {b16['kt']['force_ms'] * 1e3:.1f} us with an event "
            f"pair per launch; loss {b16['loss']:.6e}")
        bf16_bytes = (edges * (d * 2 + 4) + (g.n + 1) * 8 + 2 * g.n * d * 4 + g.n * d * 2
                      + batches * ns * d * 2) / batches
        out["bf16_form"] = {
            "value": b16["units"] / b16["dt"],
            "ms_per_step": b16["dt"] / max(args.steps, 1) * 1e3,
            "speedup_vs_fp32": f32["dt"] / b16["dt"],
            "force_avg_launch_ms": kb["force_ms"],
            "force_avg_launch_ms_event_pair_per_launch": b16["kt"]["force_ms"],
            "achieved_GBps": bf16_bytes / (kb["force_ms"] * 1e-3) / 1e9 if kb["force_ms"] else None,
            "loss_fp64_after_timed_epochs": b16["loss"],
            "loss_rel_diff_vs_fp32": (b16["loss"] - f32["loss"]) / f32["loss"],
            "tuning": b16["tuning"],
            "tuning_source": source(args.bf16_tuning),
        }
        # R-MAT training at this rate is chaotic: the summation order alone
        # (the tiling) moves the final loss. An FP32 run with the BF16 run's
        # tiling gives the like-for-like comparison and the FP32 spread.
        if not args.no_same_tiling:
            f32b = run_form(api, g, X0, lists, api.FP32, args, b16["tuning"])
            out["bf16_form"]["loss_rel_diff_vs_fp32_same_tiling"] = (
                (b16["loss"] - f32b["loss"]) / f32b["loss"])
            out["bf16_form"]["fp32_loss_rel_diff_between_tilings"] = (
                (f32b["loss"] - f32["loss"]) / f32["loss"])
            log(f"fp32 with the bf16 tiling: loss {f32b['loss']:.6e}")
        if args.loss_seeds > 0:
            from distembed_amd import bench_forms as bf
            out["bf16_form"]["loss_over_seeds"] = bf.loss_over_seeds(
                g, lists, d, B, ns, lr, args.steps, f32["tuning"], b16["tuning"], args.loss_seeds)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(g, lists, d, B, ns, lr, args.cpu_budget,
                                           args.cpu_epochs)
    del g
    from distembed_amd import bench_forms as bf
    if not args.no_precision:
        out["precision"] = bf.precision(args.steps, args.warmup)
    if not args.no_serving:
        sv = bf.serving((4.0, 8.0, 16.0), 48, 50, budget_gib=8.0)
        out["serving"] = {k: sv[k] for k in ("epochs", "batch", "ns", "budget_gib")}
        out["serving"]["levels"] = [{k: lv[k] for k in ("offered_qps", "jobs", "achieved_qps",
                                                        "p50_ms", "p99_ms", "units_per_s")}
                                    for lv in sv["levels"]]
    if not args.no_batching:
        bt = bf.batching(50, count=512, budget_gib=48.0, tune_every=1)
        out["batching"] = {k: bt[k] for k in ("graphs", "vertices", "csr_entries", "d", "batch",
                                              "ns", "epochs", "units", "tuned_on",
                                              "bf16_speedup")}
        for k in ("fp32", "bf16"):
            out["batching"][k] = {f: bt[k][f] for f in ("units_per_s", "exec_s", "rounds",
                                                       "tuning")}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
