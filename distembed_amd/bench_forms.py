"""Workloads of the precision, batching and serving forms (bench.py --form).

precision: R-MAT scale 20 and 22, d=128, FP32 and BF16-shadow gathers,
  5 and 10 negatives: epoch time, BF16 speed-up, final loss of each (also
  against an FP32 run with the BF16 run's tiling) and the per-step gradient
  error of the BF16 form against the FP32 kernel.
batching: 512 independent SBM graphs (n log-uniform in [1e3, 2e5], 4-32
  blocks, average degree uniform in [8, 40] with 90 % of it inside the
  block), d=64, B=1024, 5 negatives, 50 epochs, FP32 and BF16; all graphs
  resident, advanced together round by round, the rounds captured in HIP
  graphs (fe_engine_run_graph), tiles chosen on the mix itself by the
  library's engine autotuner (fe_engine_autotune).
serving: Poisson arrivals of R-MAT jobs (edge count log-uniform in
  [2e4, 2e7], d from {32, 64, 128}), 50 epochs, 5 negatives, admitted under
  an HBM budget and interleaved round by round (fe_engine_run); submission
  and collection run on their own threads. Job latency is arrival -> result
  on the host, reported as p50/p99 with the aggregate edges+negatives/s, at
  three offered loads.
"""
from __future__ import annotations

import math
import queue
import sys
import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import api


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def _units(g: api.Graph, B: int, ns: int, epochs: int, quirks: int = 3) -> int:
    """Edge plus negative interactions of `epochs` epochs (reference loop)."""
    n = g.n
    edges = int((g.col_idx < n - 1).sum()) if quirks & 2 else g.nnz
    batches = -(-n // B)
    last = n - B * (batches - 1)
    first = (B * (batches - 1) + last) * ns
    later = (batches * last if (quirks & 1) else B * (batches - 1) + last) * ns
    return epochs * edges + (first + (epochs - 1) * later if epochs else 0)


# ---- precision form --------------------------------------------------------

def _train_timed(g, X0, lists, gather, B, ns, lr, steps, warmup, tuning=None):
    """Autotune (unless a tuning is given), warm up, time `steps` epochs from
    X0; returns ms/epoch, units/s, the FP64 loss at the end and the tuning."""
    with api.ForceEmbed(g.n, X0.shape[1], gather=gather) as ctx:
        ctx.upload_csr(g)
        ctx.upload_x(X0)
        ctx.set_negatives(lists)
        ctx.configure(B, lr, 5.0)
        if tuning is None:
            ctx.autotune(epochs=2, search_chunking=True)
        else:
            ctx.set_tuning(**tuning["tuning"])
            ctx.set_batch_tunings(tuning["per_batch"])
        tun = {"tuning": ctx.get_tuning(),
               "per_batch": [[t[k] for k in ("unroll", "waves_per_block", "items_per_wave",
                                              "chunk", "max_chunks", "light", "slice")]
                             for t in ctx.batch_tunings()]}
        if warmup:
            ctx.upload_x(X0)
            ctx.train(warmup, B, ns, lr)
        ctx.upload_x(X0)
        ctx.synchronize()
        t1 = time.perf_counter()
        ctx.train(steps, B, ns, lr)
        ctx.synchronize()
        dt = time.perf_counter() - t1
        a, r = ctx.loss(0, 5)
        e, n0, n1 = ctx.epoch_units(ns)
    units = steps * e + n0 + (steps - 1) * n1
    return {"ms_per_epoch": dt / steps * 1e3, "units_per_s": units / dt, "loss": a + r,
            "tuning": tun}


def loss_over_seeds(g, lists, d, B, ns, lr, steps, tun32, tun16, seeds):
    """Final FP64 loss of the FP32 and the BF16 form after `steps` epochs
    from each of several initial embeddings (seeds 1..seeds, same graph and
    lists, each form with its tiling: {"tuning": ..., "per_batch": ...}).
    One run's final loss is chaotic (the summation order alone moves it by
    about a percent); the mean of the per-seed paired differences, with its
    standard error, shows whether the BF16 shadow shifts the loss."""
    ctxs = {}
    rows = []
    try:
        for gather, tun in ((api.FP32, tun32), (api.BF16, tun16)):
            ctx = api.ForceEmbed(g.n, d, gather=gather)
            ctxs[gather] = ctx
            ctx.upload_csr(g)
            ctx.set_negatives(lists)
            ctx.configure(B, lr, 5.0)
            if tun:
                ctx.set_tuning(**tun.get("tuning", {}))
        for s in range(1, seeds + 1):
            x0 = api.init_uniform(g.n, d, -0.5, 0.5, seed=s)
            row = {"seed": s}
            for gather, tun in ((api.FP32, tun32), (api.BF16, tun16)):
                ctx = ctxs[gather]
                ctx.upload_x(x0)
                if s == 1 and tun and tun.get("per_batch"):
                    ctx.set_batch_tunings(tun["per_batch"])
                ctx.train(steps, B, ns, lr)
                a, r = ctx.loss(0, ns)
                row["bf16" if gather == api.BF16 else "fp32"] = a + r
            del x0
            row["rel_diff"] = (row["bf16"] - row["fp32"]) / row["fp32"]
            rows.append(row)
            log(f"loss seed {s}: fp32 {row['fp32']:.6e} bf16 {row['bf16']:.6e} "
                f"({row['rel_diff'] * 100:+.2f}%)")
    finally:
        for ctx in ctxs.values():
            ctx.close()
    diffs = np.array([r["rel_diff"] for r in rows])
    f32 = np.array([r["fp32"] for r in rows])
    b16 = np.array([r["bf16"] for r in rows])
    return {
        "seeds": seeds, "epochs": steps,
        "mean_rel_diff": float(diffs.mean()),
        "stderr_rel_diff": float(diffs.std(ddof=1) / np.sqrt(len(diffs))) if len(diffs) > 1 else None,
        "rel_diff_of_means": float((b16.mean() - f32.mean()) / f32.mean()),
        "fp32_loss_cv_over_seeds": float(f32.std(ddof=1) / f32.mean()) if len(f32) > 1 else None,
        "per_seed": rows,
    }


def _gradient_error(g, X0, lists, B, ns, lr, batches):
    """||delta_bf16 - delta_fp32|| / ||delta_fp32|| of single steps on the
    same X and lists (the BF16 form against the plain FP32 kernel)."""
    out = []
    ctxs = []
    try:
        for gather in (api.FP32, api.BF16):
            ctx = api.ForceEmbed(g.n, X0.shape[1], gather=gather)
            ctxs.append(ctx)
            ctx.upload_csr(g)
            ctx.upload_x(X0)
            ctx.set_negatives(lists)
            ctx.configure(B, lr, 5.0)
        for (e, j) in batches:
            rows = min((j + 1) * B, g.n) - j * B
            d32 = ctxs[0].step_delta(e, j, rows).astype(np.float64)
            d16 = ctxs[1].step_delta(e, j, rows).astype(np.float64)
            out.append({"epoch": e, "batch": j,
                        "rel_err": float(np.linalg.norm(d16 - d32) / np.linalg.norm(d32))})
    finally:
        for ctx in ctxs:
            ctx.close()
    return out


def precision(steps: int, warmup: int, cases=((20, 5), (20, 10), (22, 5), (22, 10)),
              seed_cases=((22, 5),), seeds: int = 4) -> dict:
    """R-MAT at each (scale, ns) of `cases`, d=128, B=16384: FP32 and BF16
    epoch times (each autotuned), the BF16 speed-up, the final FP64 loss of
    both and of an FP32 run with the BF16 run's tiling (the tiling alone,
    i.e. the summation order, moves the final loss of these chaotic runs),
    and the per-step gradient error of the BF16 form against the FP32 kernel
    on the first, second and last batch. For the cases in `seed_cases`, the
    loss difference over `seeds` initial embeddings as well
    (loss_over_seeds)."""
    out = []
    graphs = {}
    d, B, lr = 128, 16384, 0.02
    for scale, ns in cases:
        if scale not in graphs:
            graphs.clear()
            t0 = time.perf_counter()
            graphs[scale] = api.Graph.rmat(scale, 16, 0.57, 0.19, 0.19, seed=0)
            log(f"precision: R-MAT {scale}: n={graphs[scale].n} nnz={graphs[scale].nnz} "
                f"({time.perf_counter() - t0:.1f} s)")
        g = graphs[scale]
        X0 = api.init_uniform(g.n, d, -0.5, 0.5, seed=0)
        batches = -(-g.n // B)
        lists = api.reference_negatives(g.n, steps + warmup + batches, ns)[0]
        row = {"graph": f"rmat{scale}", "n": g.n, "nnz": g.nnz, "ns": ns}
        row["fp32"] = _train_timed(g, X0, lists, api.FP32, B, ns, lr, steps, warmup)
        row["bf16"] = _train_timed(g, X0, lists, api.BF16, B, ns, lr, steps, warmup)
        same = _train_timed(g, X0, lists, api.FP32, B, ns, lr, steps, warmup,
                            tuning=row["bf16"]["tuning"])
        if (scale, ns) in seed_cases and seeds > 0:
            row["loss_over_seeds"] = loss_over_seeds(g, lists, d, B, ns, lr, steps,
                                                     row["fp32"]["tuning"], row["bf16"]["tuning"],
                                                     seeds)
        for k in ("fp32", "bf16"):
            row[k].pop("tuning")
        row["bf16_speedup"] = row["fp32"]["ms_per_epoch"] / row["bf16"]["ms_per_epoch"]
        row["bf16_loss_rel_diff_vs_fp32"] = (row["bf16"]["loss"] - row["fp32"]["loss"]) / \
            row["fp32"]["loss"]
        row["bf16_loss_rel_diff_vs_fp32_same_tiling"] = (row["bf16"]["loss"] - same["loss"]) / \
            same["loss"]
        row["fp32_loss_rel_diff_between_tilings"] = (same["loss"] - row["fp32"]["loss"]) / \
            row["fp32"]["loss"]
        row["bf16_gradient_error_vs_fp32"] = _gradient_error(
            g, X0, lists, B, ns, lr, [(0, 0), (0, 1), (0, batches - 1)])
        log(f"precision: rmat{scale} ns={ns}: fp32 {row['fp32']['ms_per_epoch']:.3f} ms, bf16 "
            f"{row['bf16']['ms_per_epoch']:.3f} ms ({row['bf16_speedup']:.3f}x), loss diff "
            f"{row['bf16_loss_rel_diff_vs_fp32']:+.2%} ({row['bf16_loss_rel_diff_vs_fp32_same_tiling']:+.2%}"
            f" same tiling)")
        out.append(row)
    return {"form": "precision", "epochs_timed": steps, "d": d, "batch": B, "lr": lr,
            "results": out}


# ---- batching form ---------------------------------------------------------

def sbm_pool(count: int, seed: int = 0):
    """`count` SBM graphs: n log-uniform in [1e3, 2e5], blocks in [4, 32],
    average degree uniform in [8, 40], 90 % of it inside the block."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        n = int(round(math.exp(rng.uniform(math.log(1e3), math.log(2e5)))))
        blocks = int(rng.integers(4, 33))
        deg = float(rng.uniform(8, 40))
        nb = n / blocks
        p_in = min(1.0, 0.9 * deg / max(nb - 1, 1))
        p_out = min(1.0, 0.1 * deg / max(n - nb, 1))
        specs.append((n, blocks, p_in, p_out, 1000 + i))
    with ThreadPoolExecutor(8) as ex:
        return list(ex.map(lambda s: api.Graph.sbm(*s), specs))


def tune_engine(graphs, X0s, B: int, epochs: int, ns: int, gather: int, budget_gib: float):
    """The library's engine autotuner (fe_engine_autotune) on the given jobs:
    each candidate tiling runs them through the captured-graph schedule and
    the fastest replay wins."""
    with api.Engine(int(budget_gib * 2 ** 30)) as eng:
        return eng.autotune([(g, x, B, dict(ns=ns, gather=gather)) for g, x in zip(graphs, X0s)],
                            epochs)


def batching(epochs: int, count: int = 512, d: int = 64, B: int = 1024, ns: int = 5,
             budget_gib: float = 96.0, tune_every: int = 1) -> dict:
    t0 = time.perf_counter()
    graphs = sbm_pool(count)
    n_tot = sum(g.n for g in graphs)
    nnz_tot = sum(g.nnz for g in graphs)
    log(f"batching: {count} SBM graphs, {n_tot} vertices, {nnz_tot} CSR entries "
        f"({time.perf_counter() - t0:.1f} s)")
    X0s = [api.init_uniform(g.n, d, -0.5, 0.5, seed=i) for i, g in enumerate(graphs)]
    units = sum(_units(g, B, ns, epochs) for g in graphs)
    # tiles for this mix: searched on every tune_every-th graph for 2 epochs (with
    # every graph, the rounds of 2 epochs hold the same mix of active jobs as
    # those of 50, in proportion: a job is active for epochs x its batches)
    res = {"form": "batching", "graphs": count, "vertices": n_tot, "csr_entries": nnz_tot,
           "d": d, "batch": B, "ns": ns, "epochs": epochs, "units": units,
           "tuned_on": ("every graph" if tune_every == 1 else f"every {tune_every}th graph")
           + ", 2 epochs"}
    for name, gather in (("fp32", api.FP32), ("bf16", api.BF16)):
        t0 = time.perf_counter()
        best, best_ms, table = tune_engine(graphs[::tune_every], X0s[::tune_every], B,
                                           min(epochs, 2), ns, gather, budget_gib)
        log(f"batching {name}: tiles {best} ({best_ms:.1f} ms on the sample; {len(table)} "
            f"candidates in {time.perf_counter() - t0:.1f} s)")
        with api.Engine(int(budget_gib * 2 ** 30)) as eng:
            eng.set_tuning(**best)
            t1 = time.perf_counter()
            ids = [eng.submit(g, x, B, epochs, ns=ns, gather=gather) for g, x in zip(graphs, X0s)]
            t_sub = time.perf_counter() - t1
            peak = eng.stats()["bytes_in_use"]
            t2 = time.perf_counter()
            eng.run_graph()
            wall = time.perf_counter() - t2
            gt = eng.graph_timing()
            xs = [eng.collect(i) for i in ids[:4]]
            assert all(np.isfinite(x).all() for x in xs)
        res[name] = {"units_per_s": units / (gt["exec_ms"] * 1e-3), "exec_s": gt["exec_ms"] * 1e-3,
                     "prepare_s": gt["prepare_ms"] * 1e-3, "wall_s": wall, "submit_s": t_sub,
                     "rounds": gt["rounds"], "hbm_bytes_resident": peak, "tuning": best}
        log(f"batching {name}: {units / (gt['exec_ms'] * 1e-3) / 1e9:.2f} G units/s "
            f"(exec {gt['exec_ms'] / 1e3:.2f} s, prepare {gt['prepare_ms'] / 1e3:.2f} s, "
            f"{gt['rounds']} rounds, {peak / 2 ** 30:.1f} GiB resident)")
    res["bf16_speedup"] = res["fp32"]["exec_s"] / res["bf16"]["exec_s"]
    return res


# ---- serving form ----------------------------------------------------------

def rmat_job_pool(count: int, seed: int = 1):
    """R-MAT jobs: generated edge count log-uniform in [2e4, 2e7] (edge factor
    16, scale from the edge count), d from {32, 64, 128}."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        m = math.exp(rng.uniform(math.log(2e4), math.log(2e7)))
        scale = max(6, int(round(math.log2(m / 16))))
        ef = max(1, int(round(m / 2 ** scale)))
        d = int(rng.choice([32, 64, 128]))
        specs.append((scale, ef, d, 2000 + i))
    with ThreadPoolExecutor(8) as ex:
        graphs = list(ex.map(lambda s: api.Graph.rmat(s[0], s[1], 0.57, 0.19, 0.19, seed=s[3]),
                             specs))
    return [(g, s[2]) for g, s in zip(graphs, specs)]


def serving(qps_levels=(4.0, 8.0, 16.0), jobs_per_level: int = 48, epochs: int = 50,
            B: int = 16384, ns: int = 5, budget_gib: float = 32.0,
            rounds_per_call: int = 16) -> dict:
    """One thread submits jobs as they arrive (retrying while the budget is
    full), one collects finished jobs, and the main thread runs rounds: a
    job's upload and download overlap the other jobs' rounds."""
    pool = rmat_job_pool(jobs_per_level)
    X0s = [api.init_uniform(g.n, d, -0.5, 0.5, seed=i) for i, (g, d) in enumerate(pool)]
    units = [_units(g, min(B, g.n), ns, epochs) for g, _ in pool]
    log(f"serving: {len(pool)} jobs, {sum(g.nnz for g, _ in pool)} CSR entries, "
        f"{sum(units) / 1e9:.2f} G units")
    levels = []
    for qps in qps_levels:
        rng = np.random.default_rng(int(qps * 1000))
        arrivals = np.cumsum(rng.exponential(1.0 / qps, len(pool)))
        arrivals -= arrivals[0]
        lat = [None] * len(pool)
        with api.Engine(int(budget_gib * 2 ** 30)) as eng:
            lock = threading.Lock()
            freed = threading.Condition()
            running = {}  # job id -> pool index
            all_submitted = threading.Event()
            to_collect: queue.Queue = queue.Queue()
            errors = []
            t0 = time.perf_counter()

            def submitter():
                try:
                    for i, (g, d) in enumerate(pool):
                        wait = arrivals[i] - (time.perf_counter() - t0)
                        if wait > 0:
                            time.sleep(wait)
                        while True:
                            try:
                                jid = eng.submit(g, X0s[i], min(B, g.n), epochs, ns=ns)
                                break
                            except api.FeError as e:
                                if e.status != 4:  # not over budget
                                    raise
                                with freed:
                                    freed.wait(0.002)
                        with lock:
                            running[jid] = i
                except Exception as e:  # reported by the main thread
                    errors.append(e)
                finally:
                    all_submitted.set()

            def collector():
                while True:
                    item = to_collect.get()
                    if item is None:
                        return
                    jid, i = item
                    try:
                        eng.collect(jid)
                    except Exception as e:
                        errors.append(e)
                    lat[i] = time.perf_counter() - t0 - arrivals[i]
                    with freed:
                        freed.notify_all()

            threads = [threading.Thread(target=submitter), threading.Thread(target=collector)]
            for t in threads:
                t.start()
            done_units = 0
            while not errors:
                with lock:
                    ids = list(running.items())
                if not ids:
                    if all_submitted.is_set():
                        with lock:
                            if not running:
                                break
                    time.sleep(0.0005)
                    continue
                left = eng.run(rounds_per_call)
                finished = [(jid, i) for jid, i in ids if eng.done(jid)]
                for jid, i in finished:
                    with lock:
                        running.pop(jid)
                    done_units += units[i]
                    to_collect.put((jid, i))
                if left == 0 and not finished:
                    time.sleep(0.0002)  # every round is enqueued: wait for the device
            to_collect.put(None)
            for t in threads:
                t.join()
            if errors:
                raise errors[0]
            span = time.perf_counter() - t0
        lat_ms = np.array(lat) * 1e3
        levels.append({"offered_qps": qps, "jobs": len(pool), "span_s": span,
                       "achieved_qps": len(pool) / span,
                       "p50_ms": float(np.percentile(lat_ms, 50)),
                       "p99_ms": float(np.percentile(lat_ms, 99)),
                       "mean_ms": float(lat_ms.mean()),
                       "units_per_s": done_units / span})
        log(f"serving qps={qps}: p50 {levels[-1]['p50_ms']:.1f} ms, p99 {levels[-1]['p99_ms']:.1f} ms, "
            f"{done_units / span / 1e9:.2f} G units/s over {span:.1f} s")
    return {"form": "serving", "epochs": epochs, "batch": B, "ns": ns, "budget_gib": budget_gib,
            "rounds_per_call": rounds_per_call, "levels": levels}
