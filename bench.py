#!/usr/bin/env python3
"""Throughput of the force-model embedding hot path on one MI355X.

Workload: R-MAT scale 20 (edge factor 16, a=0.57, b=c=0.19; symmetrised,
duplicates and self-loops removed, natural vertex order), d=128, batch 16384,
5 reference negatives per batch, lr 0.02 -- the one-GPU configuration the
metric is quoted on. A step is one epoch: the 64 dependent minibatches of
algo_force2_vec_ns over the whole graph, replayed as one captured HIP graph.
Inputs are resident in HBM before the timed region starts.

value = (processed CSR entries + (row, negative) pairs) / second.
roofline: the force kernel's algorithmic bytes per launch (its share of the
per-epoch byte count: gathered rows d*4+4 per processed entry, own row read
and new row written d*4 each, the batch's negative rows) divided by its
average launch duration from HIP events recorded on its own stream over the
timed region. cpu_baseline: the oracle's FP64 restatement of the reference
loop (OpenMP, schedule(static)) on a bounded sample of the same workload.
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


def cpu_baseline(g, lists, d, B, ns, lr, budget_s: float, max_epochs: int):
    sys.path.insert(0, str(ROOT / "tests"))
    import _oracle as orc  # the checker, timed as the CPU baseline
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
    ap.add_argument("--gather", choices=["fp32", "bf16"], default="fp32")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU baseline")
    ap.add_argument("--cpu-epochs", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--tuning", type=str, default="", help="json dict of fe_tuning fields")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the halo exchange across GPUs is a design item); use --gpus 1", file=sys.stderr)
        return 2

    from distembed_amd import _lib, api
    if _lib.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        return 3

    t0 = time.perf_counter()
    g = api.Graph.rmat(args.scale, 16, 0.57, 0.19, 0.19, seed=0)
    log(f"graph: n={g.n} nnz={g.nnz} ({time.perf_counter() - t0:.1f} s)")
    d, B, ns, lr = args.d, args.batch, args.ns, args.lr
    batches = -(-g.n // B)
    X0 = api.init_uniform(g.n, d, -0.5, 0.5, seed=0)
    epochs_needed = max(args.steps, args.warmup, args.cpu_epochs)
    lists, hits = api.reference_negatives(g.n, epochs_needed + batches - 1, ns)
    gather = api.BF16 if args.gather == "bf16" else api.FP32
    ctx = api.ForceEmbed(g.n, d, gather=gather)
    ctx.upload_csr(g)
    ctx.upload_x(X0)
    ctx.set_negatives(lists)
    if args.tuning:
        ctx.set_tuning(**json.loads(args.tuning))
    ctx.configure(B, lr, 5.0)
    edges, n0, n1 = ctx.epoch_units(ns)
    use_graph = not args.no_graph

    log(f"warmup: {args.warmup} epochs")
    if args.warmup:
        ctx.train(args.warmup, B, ns, lr, use_graph=use_graph)
    ctx.synchronize()

    # timed region: K epochs, per-launch HIP events on the context's stream
    ctx.set_timing(True)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.train(args.steps, B, ns, lr, use_graph=use_graph)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    kt = ctx.kernel_times()
    ctx.set_timing(False)
    log(f"timed: {args.steps} epochs in {dt:.3f} s; force {kt['force_ms'] * 1e3:.1f} us, "
        f"fixup {kt['fixup_ms'] * 1e3:.1f} us, apply {kt['apply_ms'] * 1e3:.1f} us per batch "
        f"({kt['launches']} launches)")
    units = args.steps * edges + (n0 + (args.steps - 1) * n1 if args.steps else 0)
    value = units / dt

    # untimed check of the result: finite, FP64 loss reduction on the device
    attr, rep = ctx.loss(0, ns)
    ctx.close()

    s_g = 2 if gather == api.BF16 else 4
    per_epoch_bytes = (edges * (d * s_g + 4) + (g.n + 1) * 8 + 2 * g.n * d * 4
                       + batches * ns * d * s_g)
    bytes_per_launch = per_epoch_bytes / batches
    force_s = kt["force_ms"] * 1e-3
    achieved = bytes_per_launch / force_s / 1e9 if force_s > 0 else None
    traffic = traffic_from_profile(f"force_kernel_rmat20_{args.gather}_pmc.txt")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "edges+negatives/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / max(args.steps, 1) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if gather == api.FP32 else "bf16-gather/fp32-accumulate",
        "data": "synthetic",
        "config": {
            "workload": f"rmat{args.scale}_d{d}_{args.gather}",
            "graph": f"R-MAT scale {args.scale}, edge factor 16, a=0.57 b=c=0.19, seed 0, "
                     "symmetrised, deduplicated, no self-loops",
            "n": g.n, "nnz": g.nnz, "d": d, "batch": B, "ns": ns, "lr": lr,
            "epochs_timed": args.steps, "step": "one epoch (64 dependent minibatches)",
            "graph_replay": use_graph, "negative_draws_equal_to_n": hits,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBPS) if achieved else None,
            "traffic": traffic,
            "kernel": "force_kernel",
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "avg_launch_ms": kt["force_ms"],
        },
        "epoch_kernel_ms": {"force": kt["force_ms"] * batches, "fixup": kt["fixup_ms"] * batches,
                            "apply": kt["apply_ms"] * batches},
        "loss_fp64": {"attractive": attr, "repulsive_eval": rep},
    }
    if not args.no_cpu:
        X_cpu_lists = lists
        out["cpu_baseline"] = cpu_baseline(g, X_cpu_lists, d, B, ns, lr, args.cpu_budget,
                                           args.cpu_epochs)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
