"""Multi-job engine (batching and serving forms) on the GPU.

Each job's result must be bitwise the same whether it runs alone, mixed
with other jobs in shared rounds, through the captured-graph schedule, or
as a single-graph context (fe_train): the work items and summation order of
a job are fixed at submission and no job touches another's memory.
"""
import queue
import threading
import time

import numpy as np
import pytest

import _oracle as orc
from distembed_amd import api

pytestmark = pytest.mark.gpu

GiB = 1 << 30


def _jobs():
    out = []
    specs = [  # (graph, d, gather, batch, epochs)
        (api.Graph.rmat(12, 16, seed=1), 128, api.FP32, 512, 4),
        (api.Graph.sbm(3000, 12, 0.01, 0.0005, seed=2), 64, api.BF16, 1024, 6),
        (api.Graph.sbm(1200, 4, 0.02, 0.001, seed=3), 32, api.FP32, 256, 3),
        (api.Graph.rmat(11, 8, seed=4), 64, api.FP32, 333, 5),
    ]
    for k, (g, d, gather, B, T) in enumerate(specs):
        x0 = api.init_uniform(g.n, d, seed=10 + k)
        out.append(dict(g=g, x0=x0, batch_size=B, iterations=T, gather=gather))
    return out


def _submit(e, spec):
    return e.submit(spec["g"], spec["x0"], spec["batch_size"], spec["iterations"],
                    gather=spec["gather"], ns=5, lr=0.02)


def _run_all(e, ids):
    while e.run(16) > 0:
        pass
    return [e.collect(i) for i in ids]


def test_alone_equals_mixed_and_graph_schedule():
    specs = _jobs()
    alone = []
    for s in specs:
        with api.Engine(2 * GiB) as e:
            alone.append(_run_all(e, [_submit(e, s)])[0])
    with api.Engine(2 * GiB) as e:
        ids = [_submit(e, s) for s in specs]
        mixed = _run_all(e, ids)
    with api.Engine(2 * GiB) as e:
        ids = [_submit(e, s) for s in specs]
        e.run_graph()
        assert all(e.done(i) for i in ids)
        graphed = [e.collect(i) for i in ids]
    for a, m, g in zip(alone, mixed, graphed):
        assert np.isfinite(a).all()
        np.testing.assert_array_equal(a, m)
        np.testing.assert_array_equal(a, g)


def test_every_round_size_gives_the_alone_bits():
    """Eleven jobs of different lengths: rounds of more than eight jobs
    (descriptors copied with the block-to-step table), of two to eight
    (passed as kernel arguments, blocks search the prefix) and of one (the
    single-step kernel) all give each job its alone result. Some widths are
    held padded on the device (3, 100, 12, 2)."""
    specs = _jobs()
    for k in range(7):
        g = api.Graph.sbm(600 + 300 * k, 3 + k, 0.03, 0.002, seed=20 + k)
        d = (32, 3, 100, 64, 128, 12, 2)[k]
        specs.append(dict(g=g, x0=api.init_uniform(g.n, d, seed=30 + k), batch_size=256,
                          iterations=2 + k, gather=api.BF16 if k % 2 else api.FP32))
    alone = []
    for s in specs:
        with api.Engine(2 * GiB) as e:
            alone.append(_run_all(e, [_submit(e, s)])[0])
    with api.Engine(2 * GiB) as e:
        ids = [_submit(e, s) for s in specs]
        mixed = _run_all(e, ids)
    with api.Engine(2 * GiB) as e:
        ids = [_submit(e, s) for s in specs]
        e.run_graph()
        graphed = [e.collect(i) for i in ids]
    for a, m, g in zip(alone, mixed, graphed):
        assert np.isfinite(a).all()
        np.testing.assert_array_equal(a, m)
        np.testing.assert_array_equal(a, g)


def test_engine_batch_size_extremes():
    """A job of batch size 1 (one row per round) beside a BF16 job whose
    batch size exceeds its row count (one batch of all rows): each equals
    its alone run, mixed and through the captured-graph schedule, and the
    single-graph context."""
    g1 = api.Graph.sbm(200, 2, 0.05, 0.01, seed=40)
    g2 = api.Graph.sbm(300, 3, 0.05, 0.01, seed=41)
    specs = [dict(g=g1, x0=api.init_uniform(g1.n, 32, seed=42), batch_size=1, iterations=2,
                  gather=api.FP32),
             dict(g=g2, x0=api.init_uniform(g2.n, 64, seed=43), batch_size=4096, iterations=3,
                  gather=api.BF16)]
    alone = []
    for s in specs:
        with api.Engine(1 * GiB) as e:
            alone.append(_run_all(e, [_submit(e, s)])[0])
    with api.Engine(1 * GiB) as e:
        ids = [_submit(e, s) for s in specs]
        mixed = _run_all(e, ids)
    with api.Engine(1 * GiB) as e:
        ids = [_submit(e, s) for s in specs]
        e.run_graph()
        graphed = [e.collect(i) for i in ids]
    for s, a, m, gr in zip(specs, alone, mixed, graphed):
        np.testing.assert_array_equal(a, m)
        np.testing.assert_array_equal(a, gr)
        g = s["g"]
        with api.ForceEmbed(g.n, s["x0"].shape[1], gather=s["gather"]) as ctx:
            ctx.upload_csr(g)
            ctx.upload_x(s["x0"])
            ctx.train(s["iterations"], s["batch_size"], 5, 0.02)
            np.testing.assert_array_equal(a, ctx.download_x())


def test_engine_job_equals_single_graph_context():
    s = _jobs()[0]
    with api.Engine(2 * GiB) as e:
        xe = _run_all(e, [_submit(e, s)])[0]
    g = s["g"]
    with api.ForceEmbed(g.n, s["x0"].shape[1], gather=s["gather"]) as ctx:
        ctx.upload_csr(g)
        ctx.upload_x(s["x0"])
        ctx.train(s["iterations"], s["batch_size"], 5, 0.02)
        xc = ctx.download_x()
    np.testing.assert_array_equal(xe, xc)


def test_every_job_of_mixed_rounds_matches_context_and_oracle():
    """The four job kinds (FP32 d=128, BF16 d=64, FP32 d=32, B=333 with the
    multiply-high slot division) share rounds whose launches mix FP32 and
    BF16 gathers and lane counts (the kernel's run-time dtype and width
    form). Each result equals the single-graph context bit for bit and tracks
    the oracle's run: FP32 within 1e-4 of max|X| everywhere and 1e-5
    normwise plus twice the FP32 oracle's own drift from the FP64 oracle;
    the BF16 form against the oracle's BF16 form within 1e-5 normwise and
    1e-3 of max|X|."""
    specs = _jobs()
    with api.Engine(2 * GiB) as e:
        ids = [_submit(e, s) for s in specs]
        mixed = _run_all(e, ids)
    for k, (s, xe) in enumerate(zip(specs, mixed)):
        g, x0, B, T = s["g"], s["x0"], s["batch_size"], s["iterations"]
        with api.ForceEmbed(g.n, x0.shape[1], gather=s["gather"]) as ctx:
            ctx.upload_csr(g)
            ctx.upload_x(x0)
            ctx.train(T, B, 5, 0.02)
            np.testing.assert_array_equal(xe, ctx.download_x(), err_msg=f"job {k}")
        lists = api.reference_negatives(g.n, T + -(-g.n // B) - 1, 5)[0]
        xr = x0.copy()
        orc.train(g, xr, B, T, lists, bf16=(s["gather"] == api.BF16))
        diff = xe.astype(np.float64) - xr
        if s["gather"] == api.BF16:
            assert np.linalg.norm(diff) <= 1e-5 * np.linalg.norm(xr), k
            assert np.abs(diff).max() <= 1e-3 * np.abs(xr).max(), k
        else:
            x64 = x0.astype(np.float64)
            orc.train(g, x64, B, T, lists)
            assert np.abs(diff).max() <= 1e-4 * np.abs(xr).max(), k
            assert np.linalg.norm(diff) <= 1e-5 * np.linalg.norm(xr) + 2 * np.linalg.norm(xr - x64), k


def test_admission_by_budget_and_reuse():
    g = api.Graph.rmat(14, 16, seed=7)
    x0 = api.init_uniform(g.n, 128, seed=1)
    with api.Engine(1 << 20) as probe:
        need = probe.job_bytes(g, x0, 2048, 2)
    assert need > 8 * g.n * 128 // 2
    with api.Engine(need * 3 // 2) as e:  # room for one copy, not two
        a = e.submit(g, x0, 2048, 2)
        with pytest.raises(api.FeError) as err:
            e.submit(g, x0, 2048, 2)  # the second copy does not fit
        assert err.value.status == 4
        used = e.stats()["bytes_in_use"]
        assert used >= need * 0.5
        while e.run(8) > 0:
            pass
        xa = e.collect(a)
        assert e.stats()["bytes_in_use"] == 0
        b = e.submit(g, x0, 2048, 2)  # memory returned to the budget
        while e.run(8) > 0:
            pass
        np.testing.assert_array_equal(e.collect(b), xa)
        big = api.Graph.rmat(17, 16, seed=7)
        with pytest.raises(api.FeError) as err:
            e.submit(big, api.init_uniform(big.n, 128), 4096, 1)
        assert err.value.status == 4


def test_bad_job_is_rejected_and_engine_stays_usable():
    g = api.Graph.sbm(1200, 4, 0.02, 0.001, seed=3)
    x0 = api.init_uniform(g.n, 32, seed=1)
    rp = g.row_ptr.copy()
    rp[5], rp[6] = rp[6], rp[5] + 1  # not monotone
    bad = api.Graph(g.n, rp, g.col_idx)
    with api.Engine(1 * GiB) as e:
        with pytest.raises(api.FeError) as err:
            e.submit(bad, x0, 256, 2)
        assert err.value.status == 1
        assert e.stats()["bytes_in_use"] == 0
        jid = e.submit(g, x0, 256, 2)
        while e.run(8) > 0:
            pass
        x = e.collect(jid)
    with api.Engine(1 * GiB) as e:
        ref = e.submit(g, x0, 256, 2)
        while e.run(8) > 0:
            pass
        np.testing.assert_array_equal(x, e.collect(ref))


def test_jobs_admitted_while_others_run():
    specs = _jobs()
    with api.Engine(2 * GiB) as e:
        first = _submit(e, specs[0])
        e.run(3)  # job 0 is mid-way when job 1 arrives
        second = _submit(e, specs[1])
        while e.run(4) > 0:
            pass
        x0, x1 = e.collect(first), e.collect(second)
    for s, x in ((specs[0], x0), (specs[1], x1)):
        with api.Engine(2 * GiB) as e:
            np.testing.assert_array_equal(_run_all(e, [_submit(e, s)])[0], x)


def test_submit_and_collect_on_other_threads():
    """fe_engine_submit and fe_engine_collect from their own threads while
    the main thread runs rounds: every result equals the job run alone."""
    specs = _jobs()
    alone = []
    for s in specs:
        with api.Engine(2 * GiB) as e:
            alone.append(_run_all(e, [_submit(e, s)])[0])
    results, errors = {}, []
    with api.Engine(2 * GiB) as e:
        ids, lock, todo = {}, threading.Lock(), queue.Queue()

        def submitter():
            try:
                for k, s in enumerate(specs):
                    jid = _submit(e, s)
                    with lock:
                        ids[jid] = k
            except Exception as ex:
                errors.append(ex)

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
        handed = set()
        deadline = time.time() + 120
        while len(handed) < len(specs) and not errors:
            assert time.time() < deadline, "jobs did not finish"
            e.run(4)
            with lock:
                cur = dict(ids)
            for jid, k in cur.items():
                if jid not in handed and e.done(jid):
                    handed.add(jid)
                    todo.put((jid, k))
        ts.join()
        todo.put(None)
        tc.join()
    assert not errors, errors
    for k, a in enumerate(alone):
        np.testing.assert_array_equal(results[k], a)


def test_engine_autotune_keeps_the_fastest_tiles():
    """fe_engine_autotune times its 72 candidate tilings (slices of 8 and
    16: every job has d >= 16) on the given jobs through the captured-graph
    schedule, keeps the fastest as the engine's tuning and leaves no job
    behind: a job run afterwards equals the same job under an engine given
    that tuning explicitly."""
    specs = _jobs()
    with api.Engine(2 * GiB) as e:
        best, best_ms, table = e.autotune(
            [(s["g"], s["x0"], s["batch_size"], dict(ns=5, gather=s["gather"])) for s in specs],
            epochs=1)
        assert len(table) == 72
        assert {t["slice"] for t, _ in table} == {8, 16}
        assert best_ms == min(ms for _, ms in table) and best_ms > 0
        assert best in [t for t, _ in table]
        assert e.stats()["bytes_in_use"] == 0
        xa = _run_all(e, [_submit(e, specs[1])])[0]
    with api.Engine(2 * GiB) as e:
        e.set_tuning(**best)
        xb = _run_all(e, [_submit(e, specs[1])])[0]
    np.testing.assert_array_equal(xa, xb)
    with api.Engine(2 * GiB) as e:
        _submit(e, specs[2])
        with pytest.raises(api.FeError) as err:  # jobs not collected yet
            e.autotune([(specs[0]["g"], specs[0]["x0"], 512, {})], epochs=1)
        assert err.value.status == 5
