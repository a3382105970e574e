"""The workloads of the precision, batching and serving forms on the GPU,
each checked against the oracle, and the reference's own CPU workload
(SBM-10K, 200 epochs) run whole.

precision: R-MAT scale 20 and 22 (a=0.57, b=c=0.19, seed 0), d=128,
  B=16384, 5 and 10 negatives. Per-batch parity of the BF16 shadow against
  the oracle at the rounded rows, per-batch parity of the FP32 path, and the
  per-step gradient error of the BF16 form against the FP64 oracle at the
  unrounded rows (the reference's own precision, cpp/core/common.h:35).
batching: the 512 SBM graphs of the batching form (n log-uniform in
  [1e3, 2e5], 4-32 blocks, average degree in [8, 40]), d=64, B=1024,
  5 negatives, FP32 and BF16, all run together through the captured-graph
  schedule (fe_engine_run_graph). Every result is finite; the smallest and
  largest graphs equal the same job alone and track the oracle.
serving: 12 R-MAT jobs drawn from the serving distribution (edge count
  log-uniform in [2e4, 2e7], one of 1.4e7, d from {32, 64, 128}), Poisson
  arrivals at 200 jobs/s, admission under an HBM budget that holds the
  largest job and the smallest one, so later arrivals wait for memory,
  submission and collection on their own threads. Every result equals the
  same job run alone bit for bit; a sample tracks the oracle.
The loop each job reproduces is cpp/algo/algo.hpp:137-192.
"""
import queue
import threading
import time

import numpy as np
import pytest

import _oracle as orc
from _parity import batch_parity, bf16_batch_parity
from distembed_amd import api
from distembed_amd import bench_forms as bf

pytestmark = pytest.mark.gpu

GiB = 1 << 30


def _track_oracle(xg, g, x0, B, T, lists, gather, what):
    """A multi-epoch run against the oracle's run on the same inputs. FP32:
    |Xg - X32| <= 1e-4 max|X32| everywhere and ||Xg - X32|| <= 1e-5 ||X32||
    + 2 ||X32 - X64|| (the FP32 oracle's own rounding drift, measured by the
    FP64 oracle). BF16 form, against the oracle's BF16 form: normwise 1e-5,
    as per batch. Elementwise the runs part where FP32 rounding (the GPU
    sums in another order) moves a master value across a BF16 rounding
    boundary: one gathered component then differs by a BF16 unit (2^-8
    relative) for the rest of the batch. Such flips are rare and their
    effects small, so at most one element in 10^5 may differ by more than
    1e-3 max|X|, and none by more than 2^-6 max|X| (four BF16 units at the
    largest magnitude)."""
    xr = x0.copy()
    orc.train(g, xr, B, T, lists, bf16=(gather == api.BF16))
    assert np.isfinite(xg).all(), what
    diff = xg.astype(np.float64) - xr
    if gather == api.BF16:
        top = np.abs(xr).max()
        assert np.linalg.norm(diff) <= 1e-5 * np.linalg.norm(xr), what
        far = int((np.abs(diff) > 1e-3 * top).sum())
        assert far <= max(1, diff.size // 100000), (what, far)
        assert np.abs(diff).max() <= 2.0 ** -6 * top, what
        return
    x64 = x0.astype(np.float64)
    orc.train(g, x64, B, T, lists)
    assert np.abs(diff).max() <= 1e-4 * np.abs(xr).max(), what
    assert np.linalg.norm(diff) <= (1e-5 * np.linalg.norm(xr)
                                    + 2 * np.linalg.norm(xr.astype(np.float64) - x64)), what


# ---- the reference's CPU workload, whole run ---------------------------------------

def test_sbm10k_full_run_tracks_oracle():
    """SBM with 10,000 vertices in 10 blocks (average intra-degree 18,
    inter-degree 2), d=128, X0 U(-0.5,0.5) seed 0, B=256 (40 batches, the
    last of 16 rows: the sticky repulsive count from epoch 1 on), 5
    reference negatives, lr 0.02, all 200 epochs. The FP32 run tracks the
    FP32 and FP64 oracles to the bars of _track_oracle: this workload is not
    chaotic, the FP32 and FP64 oracle runs end 6e-7 apart (max relative).
    The BF16 form ends within 1 % of the FP32 final loss. Its rows are
    compared with the oracle's BF16 run at the size of the BF16 effect
    itself: over 200 epochs a BF16 run moves with the rounding of its FP32
    sums (one flipped BF16 rounding changes a gathered component by 2^-9).
    The oracle's BF16 run ends 4.6e-3 (normwise) from its FP32 run, and
    5.6e-3 from itself started with 1000 components moved by one ulp. So
    the GPU's BF16 run must stay within twice the oracle's BF16-FP32
    distance normwise and four times it elementwise."""
    g = api.Graph.sbm(10_000, 10, 18 / 999, 2 / 9000, seed=0)
    d, B, T, ns = 128, 256, 200, 5
    x0 = api.init_uniform(g.n, d, -0.5, 0.5, seed=0)
    lists = api.reference_negatives(g.n, T + -(-g.n // B) - 1, ns)[0]
    out, loss = {}, {}
    for gather in (api.FP32, api.BF16):
        with api.ForceEmbed(g.n, d, gather=gather) as ctx:
            ctx.upload_csr(g)
            ctx.upload_x(x0)
            ctx.set_negatives(lists)
            ctx.configure(B, 0.02, 5.0)
            ctx.train(T, B, ns, 0.02)
            out[gather] = ctx.download_x()
            a, r = ctx.loss(0, ns)
            loss[gather] = a + r
    _track_oracle(out[api.FP32], g, x0, B, T, lists, api.FP32, "SBM-10K FP32, 200 epochs")
    x32, x16 = x0.copy(), x0.copy()
    orc.train(g, x32, B, T, lists)
    orc.train(g, x16, B, T, lists, bf16=True)
    norm = lambda a: float(np.linalg.norm(a.astype(np.float64)))
    effect, effect_max = norm(x16 - x32), float(np.abs(x16 - x32).max())
    dist, dist_max = norm(out[api.BF16] - x16), float(np.abs(out[api.BF16] - x16).max())
    print(f"  SBM-10K BF16: GPU to oracle {dist / norm(x16):.2e} normwise, {dist_max:.2e} max; "
          f"oracle BF16 to FP32 {effect / norm(x16):.2e}, {effect_max:.2e}", flush=True)
    assert np.isfinite(out[api.BF16]).all()
    assert dist <= 2.0 * effect, (dist, effect)
    assert dist_max <= 4.0 * effect_max, (dist_max, effect_max)
    rel = (loss[api.BF16] - loss[api.FP32]) / loss[api.FP32]
    print(f"  SBM-10K 200 epochs: FP32 loss {loss[api.FP32]:.6e}, BF16 {loss[api.BF16]:.6e} "
          f"({rel * 100:+.3f}%)", flush=True)
    assert abs(rel) <= 0.01, rel


# ---- precision form --------------------------------------------------------------

@pytest.fixture(scope="module")
def rmat20():
    return api.Graph.rmat(20, 16, 0.57, 0.19, 0.19, seed=0)


@pytest.fixture(scope="module")
def rmat22():
    return api.Graph.rmat(22, 16, 0.57, 0.19, 0.19, seed=0)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("ns", [5, 10])
def test_precision_rmat20_bf16_batches(rmat20, ns):
    # batch 0 holds the hub rows (5.6 M CSR entries), batch 63 the tail;
    # (1, 17) is a batch of the second epoch
    res = bf16_batch_parity(rmat20, 128, 16384, [(0, 0), (0, 1), (0, 63), (1, 17)], ns=ns,
                            seed=0, grad_err_vs_fp64=1e-2)
    for (e, j, err, ge) in res:
        print(f"rmat20 bf16 ns={ns} epoch {e} batch {j}: delta error vs rounded-row oracle "
              f"{err:.2e}, gradient error vs FP64 {ge:.2e}", flush=True)


@pytest.mark.timeout(600)
def test_precision_rmat20_fp32_ten_negatives(rmat20):
    batch_parity(rmat20, 128, 16384, [(0, 0), (2, 63)], ns=10, seed=0)


@pytest.mark.timeout(900)
def test_precision_rmat22_fp32_batches(rmat22):
    g = rmat22
    assert g.n == 1 << 22
    X = api.init_uniform(g.n, 128, -0.5, 0.5, seed=0)
    # 256 batches: batch 0 (hub rows), a middle batch and the tail
    errs = batch_parity(g, 128, 16384, [(0, 0), (1, 255)], ns=5, X=X)
    errs += batch_parity(g, 128, 16384, [(0, 128)], ns=10, X=X, apply_check=False)
    print(f"rmat22 fp32 normwise delta errors {['%.2e' % e for e in errs]}", flush=True)


@pytest.mark.timeout(900)
def test_precision_rmat22_bf16_batches(rmat22):
    g = rmat22
    X = api.init_uniform(g.n, 128, -0.5, 0.5, seed=0)
    for ns, steps in ((5, [(0, 0)]), (10, [(1, 255), (0, 64)])):
        for (e, j, err, ge) in bf16_batch_parity(g, 128, 16384, steps, ns=ns, X=X,
                                                 grad_err_vs_fp64=1e-2):
            print(f"rmat22 bf16 ns={ns} epoch {e} batch {j}: delta error {err:.2e}, "
                  f"gradient error vs FP64 {ge:.2e}", flush=True)


# ---- batching form ---------------------------------------------------------------

@pytest.fixture(scope="module")
def pool512():
    graphs = bf.sbm_pool(512)
    X0s = [api.init_uniform(g.n, 64, -0.5, 0.5, seed=i) for i, g in enumerate(graphs)]
    return graphs, X0s


@pytest.mark.timeout(900)
@pytest.mark.parametrize("gather", [api.FP32, api.BF16])
def test_batching_512_graphs(pool512, gather):
    graphs, X0s = pool512
    assert len(graphs) == 512 and sum(g.nnz for g in graphs) > 4e8
    B, T, ns = 1024, 2, 5
    order = np.argsort([g.n for g in graphs], kind="stable")
    sample = [int(order[0]), int(order[1]), int(order[-1])]
    res = {}
    with api.Engine(48 * GiB) as e:
        ids = [e.submit(g, x, B, T, ns=ns, gather=gather) for g, x in zip(graphs, X0s)]
        e.run_graph()
        t = e.graph_timing()
        assert all(e.done(i) for i in ids)
        for k, i in enumerate(ids):
            x = e.collect(i)
            assert np.isfinite(x).all(), k
            assert not np.array_equal(x, X0s[k]), k  # every job moved
            if k in sample:
                res[k] = x
    print(f"batching gather={gather}: {t['rounds']} rounds in {t['exec_ms']:.1f} ms", flush=True)
    for k in sample:
        g, x0 = graphs[k], X0s[k]
        with api.Engine(4 * GiB) as e:  # the same job alone
            i = e.submit(g, x0, B, T, ns=ns, gather=gather)
            while e.run(64) > 0:
                pass
            np.testing.assert_array_equal(e.collect(i), res[k])
        batches = -(-g.n // B)
        lists = api.reference_negatives(g.n, T + batches - 1, ns)[0]
        _track_oracle(res[k], g, x0, B, T, lists, gather, f"graph {k} (n={g.n})")


# ---- serving form ----------------------------------------------------------------

def test_serving_poisson_jobs_equal_alone():
    pool = bf.rmat_job_pool(12, seed=1)
    edges = sorted(g.nnz for g, _ in pool)
    assert edges[-1] >= 1e7 and edges[0] < 2e5
    assert {d for _, d in pool} == {32, 64, 128}
    T, ns, Bmax = 3, 5, 16384
    X0s = [api.init_uniform(g.n, d, -0.5, 0.5, seed=i) for i, (g, d) in enumerate(pool)]
    with api.Engine(1 << 20) as probe:
        need = [probe.job_bytes(g, x, min(Bmax, g.n), T) for (g, _), x in zip(pool, X0s)]
    # room for the largest job and the smallest: later arrivals wait for memory
    budget = max(need) + min(need)
    assert sum(need) > budget
    # 200 jobs/s offered: arrivals overlap the running jobs
    rng = np.random.default_rng(200)
    arrivals = np.cumsum(rng.exponential(1.0 / 200.0, len(pool)))
    arrivals -= arrivals[0]
    results, errors, waited = {}, [], []
    with api.Engine(budget) as e:
        lock = threading.Lock()
        running = {}
        all_in = threading.Event()
        todo = queue.Queue()
        t0 = time.perf_counter()

        def submitter():
            try:
                for i, (g, d) in enumerate(pool):
                    dt = arrivals[i] - (time.perf_counter() - t0)
                    if dt > 0:
                        time.sleep(dt)
                    tries = 0
                    while True:
                        try:
                            jid = e.submit(g, X0s[i], min(Bmax, g.n), T, ns=ns)
                            break
                        except api.FeError as ex:
                            if ex.status != 4:  # only over-budget is retried
                                raise
                            tries += 1
                            if tries % 5000 == 0:
                                print(f"serving: job {i} still waiting for memory", flush=True)
                            assert tries < 30000, f"job {i} was never admitted"
                            time.sleep(0.001)
                    if tries:
                        waited.append(i)
                    with lock:
                        running[jid] = i
            except Exception as ex:
                errors.append(ex)
            finally:
                all_in.set()

        def collector():
            while True:
                item = todo.get()
                if item is None:
                    return
                try:
                    results[item[1]] = e.collect(item[0])
                except Exception as ex:
                    errors.append(ex)

        ts, tc = threading.Thread(target=submitter), threading.Thread(target=collector)
        ts.start()
        tc.start()
        deadline = time.time() + 120
        while not errors:
            assert time.time() < deadline, "serving loop did not finish"
            with lock:
                cur = list(running.items())
            if not cur:
                if all_in.is_set():
                    with lock:
                        if not running:
                            break
                time.sleep(0.0005)
                continue
            e.run(16)
            for jid, i in cur:
                if e.done(jid):
                    with lock:
                        running.pop(jid)
                    todo.put((jid, i))
        ts.join()
        todo.put(None)
        tc.join()
    assert not errors, errors
    assert len(results) == len(pool)
    print(f"serving: {len(waited)} of {len(pool)} jobs waited for memory", flush=True)
    assert waited, "no arrival had to wait for memory"
    for i, ((g, d), x0) in enumerate(zip(pool, X0s)):
        B = min(Bmax, g.n)
        with api.Engine(4 * GiB) as e:
            j = e.submit(g, x0, B, T, ns=ns)
            while e.run(64) > 0:
                pass
            np.testing.assert_array_equal(e.collect(j), results[i], err_msg=f"job {i}")
    # a sample against the oracle: the two smallest jobs and a middle one
    order = np.argsort([g.nnz for g, _ in pool], kind="stable")
    for i in (int(order[0]), int(order[1]), int(order[len(order) // 2])):
        (g, d), x0 = pool[i], X0s[i]
        B = min(Bmax, g.n)
        lists = api.reference_negatives(g.n, T + -(-g.n // B) - 1, ns)[0]
        _track_oracle(results[i], g, x0, B, T, lists, api.FP32, f"job {i} (nnz={g.nnz}, d={d})")
