"""GPU parity of the force step with slices of 16 components per lane
(fe_tuning.slice = 16: a row of d components on d/16 lanes, 128/d rows per
lane group of a wave), through the C ABI, against the oracle.

The bar is the one of tests/_parity.py. A slice of 16 sums each lane's part
of |v|^2 over 16 components and reduces over half as many lanes, so its
results differ from slices of 8 at rounding level; for one slice they are
bitwise independent of the launch shape.
"""
import numpy as np
import pytest

import _oracle as orc
from _parity import (RTOL, batch_parity as _batch_parity, bf16_batch_parity,
                     clamp_fractions, make_ctx as _ctx, rows_of as _rows)
from distembed_amd import api

pytestmark = pytest.mark.gpu

S16 = dict(slice=16)


def _t(**kw):
    return dict(S16, **kw)


@pytest.mark.parametrize("d", [128, 64, 32])
def test_sbm_batches_slice16(d):
    g = api.Graph.sbm(2000, 10, 18 / 199, 2 / 1800, seed=0)
    _batch_parity(g, d, 256, [(0, 0), (0, 3), (0, 7), (1, 0), (1, 7), (4, 2)], tuning=_t())
    _batch_parity(g, d, 256, [(0, 0), (1, 7)], tuning=_t(light=32))


@pytest.mark.parametrize("d", [16, 256, 512])
def test_other_widths_slice16(d):
    g = api.Graph.sbm(700, 7, 0.05, 0.004, seed=3)
    _batch_parity(g, d, 128, [(0, 0), (2, 5)], ns=5 if d <= 128 else 3, tuning=_t())
    g = api.Graph.sbm(700, 7, 0.25, 0.004, seed=3)
    _batch_parity(g, d, 128, [(1, 2)], ns=5 if d <= 128 else 3, tuning=_t(light=32))


def test_hub_rows_split_into_chunks_slice16():
    g = api.Graph.rmat(12, 16, seed=5)
    for tun in ({"chunk": 4, "max_chunks": 8}, {"chunk": 32, "max_chunks": 256},
                {"chunk": 64, "max_chunks": 1}, {"chunk": 32, "light": -1},
                {"chunk": 64, "max_chunks": 64, "light": 16},
                {"chunk": 64, "max_chunks": 64, "light": 32}):
        _batch_parity(g, 128, 512, [(0, 0), (0, 1), (3, 7)], tuning=_t(**tun))


def test_quirks_off_no_negatives_and_empty_rows_slice16():
    g = api.Graph.sbm(1000, 4, 0.02, 0.002, seed=9)
    _batch_parity(g, 64, 300, [(0, 0), (2, 3)], quirks=0, tuning=_t())
    _batch_parity(g, 64, 300, [(0, 0), (2, 3)], ns=0, tuning=_t())
    _batch_parity(g, 64, 300, [(1, 3)], ns=10, tuning=_t())
    n = 300
    g = api.Graph.from_edges(n, np.array([0, 5]), np.array([1, 6]))
    _batch_parity(g, 64, 64, [(0, 0), (0, 4), (2, 2)], tuning=_t())
    n = 64
    g = api.Graph.from_edges(n, np.arange(n - 1), np.full(n - 1, n - 1))
    _batch_parity(g, 32, 16, [(0, 0), (0, 3)], quirks=3, tuning=_t())


@pytest.mark.parametrize("bound", [0.5, 1.0, 5.0])
def test_clamp_slice16(bound):
    # rows of small components plus two large ones: the repulsive clamp binds
    # below 5, the attractive one below 1
    g = api.Graph.sbm(2000, 10, 18 / 199, 2 / 1800, seed=0)
    d, B = 64, 256
    rng = np.random.default_rng(11)
    X = rng.uniform(-0.01, 0.01, (g.n, d)).astype(np.float32)
    k = rng.integers(0, d, (g.n, 2))
    X[np.arange(g.n)[:, None], k] = rng.uniform(-1, 1, (g.n, 2)).astype(np.float32)
    lists = api.reference_negatives(g.n, 20, 5)[0]
    steps = [(0, 0), (0, 3), (1, 7)]
    fa, fr = clamp_fractions(g, X, B, 3, lists[3], B, bound)
    if bound < 5.0:
        assert fr > 1e-3, fr
    if bound < 1.0:
        assert fa > 0.01, fa
    _batch_parity(g, d, B, steps, X=X, bound=bound, lists=lists, tuning=_t())
    _batch_parity(g, d, B, steps, X=X, bound=bound, lists=lists,
                  tuning=_t(light=0, chunk=16, max_chunks=4))


def test_non_finite_rows_slice16():
    g = api.Graph.sbm(1200, 6, 0.02, 0.001, seed=2)
    d, B = 64, 256
    X = api.init_uniform(g.n, d, seed=4)
    nan_row, inf_row, big_row = 40, 300, 700
    X[nan_row] = np.nan
    X[inf_row, 5] = np.inf
    X[big_row] = 1e20
    lists = api.reference_negatives(g.n, 12, 5)[0]
    lists[3, 1] = inf_row
    lists[4, 2] = big_row
    lists[lists == nan_row] = nan_row + 1
    ctx = _ctx(g, d, X, B, lists, tuning=_t())
    X64 = X.astype(np.float64)
    try:
        for (e, j) in ((0, 0), (0, 1), (0, 2), (1, 2), (1, 3)):
            rows = _rows(g.n, B, j)
            r0 = j * B
            dg = ctx.step_delta(e, j, rows)
            dr = orc.step(g, X, B, e, j, lists, apply=False, want_delta=True)
            dr64 = orc.step(g, X64, B, e, j, lists, apply=False, want_delta=True,
                            dtype=np.float64)
            ng, nr = X[r0:r0 + rows] + dg, X[r0:r0 + rows] + dr
            assert np.array_equal(np.isnan(ng), np.isnan(nr)), (e, j)
            assert np.array_equal(np.isposinf(ng), np.isposinf(nr)), (e, j)
            assert np.array_equal(np.isneginf(ng), np.isneginf(nr)), (e, j)
            fin = np.isfinite(nr)
            n64 = X64[r0:r0 + rows] + dr64
            ref, got = nr[fin].astype(np.float64), ng[fin].astype(np.float64)
            scale = float(np.sqrt(np.mean(ref ** 2)))
            bound = RTOL * (np.abs(ref) + scale) + 2.0 * np.abs(ref - n64[fin])
            assert np.all(np.abs(got - ref) <= bound), (e, j)
        assert np.isnan(ctx.step_delta(1, 2, B)).any()
    finally:
        ctx.close()


def test_launch_shape_does_not_change_results_slice16():
    g = api.Graph.rmat(11, 16, seed=2)
    X = api.init_uniform(g.n, 128, seed=1)
    lists = api.reference_negatives(g.n, 20, 5)[0]
    outs = []
    for tun in ({"unroll": 1, "waves_per_block": 1, "items_per_wave": 1},
                {"unroll": 4, "waves_per_block": 4, "items_per_wave": 2},
                {"unroll": 8, "waves_per_block": 8, "items_per_wave": 4},
                {"unroll": 2, "waves_per_block": 2, "items_per_wave": 16}):
        ctx = _ctx(g, 128, X, 256, lists, tuning=_t(**tun))
        ctx.train(3, 256, 5, 0.02, use_graph=False)
        outs.append(ctx.download_x())
        ctx.close()
    for o in outs[1:]:
        np.testing.assert_array_equal(o, outs[0])


def test_mixed_slices_per_batch_and_graph_replay():
    # batches alternating between slices of 8 and 16: each batch equals the
    # run whose slice is uniform at that batch; the captured epoch equals
    # stream launches and single steps bit for bit
    g = api.Graph.rmat(12, 16, seed=5)
    X = api.init_uniform(g.n, 64, seed=3)
    B = 512
    nb = -(-g.n // B)
    lists = api.reference_negatives(g.n, 20, 5)[0]
    per = [(2, 4, 1, 64, 1024, 16, 16 if b % 2 else 8) for b in range(nb)]
    ctx = _ctx(g, 64, X, B, lists)
    ctx.set_batch_tunings(per)
    assert [t["slice"] for t in ctx.batch_tunings()] == [p[6] for p in per]
    d0, d1 = ctx.step_delta(0, 0, B), ctx.step_delta(0, 1, B)
    ctx.close()
    for sl, j, dj in ((8, 0, d0), (16, 1, d1)):
        ctx = _ctx(g, 64, X, B, lists, tuning=dict(chunk=64, max_chunks=1024, light=16, slice=sl))
        np.testing.assert_array_equal(dj, ctx.step_delta(0, j, B))
        ctx.close()
    res = []
    for mode in ("graph", "stream", "steps"):
        ctx = _ctx(g, 64, X, B, lists)
        ctx.set_batch_tunings(per)
        if mode == "steps":
            for i in range(3):
                for j in range(nb):
                    ctx.step(i, j)
        else:
            ctx.train(3, B, 5, 0.02, use_graph=(mode == "graph"))
        res.append(ctx.download_x())
        ctx.close()
    np.testing.assert_array_equal(res[0], res[1])
    np.testing.assert_array_equal(res[0], res[2])


def test_training_tracks_oracle_slice16():
    g = api.Graph.sbm(1000, 10, 18 / 99, 2 / 900, seed=0)
    X = api.init_uniform(g.n, 128, seed=0)
    B, T = 256, 10
    lists = api.reference_negatives(g.n, T + 3, 5)[0]
    ctx = _ctx(g, 128, X, B, lists, tuning=_t())
    ctx.train(T, B, 5, 0.02)
    Xg = ctx.download_x()
    ctx.close()
    Xr = X.copy()
    orc.train(g, Xr, B, T, lists)
    rel = np.abs(Xg - Xr).max() / np.abs(Xr).max()
    assert rel < 1e-4, rel


def test_bf16_batches_slice16():
    g = api.Graph.rmat(12, 16, seed=5)
    for tun in (_t(), _t(chunk=32, max_chunks=64, light=0, unroll=1), _t(light=32, unroll=4)):
        bf16_batch_parity(g, 128, 512, [(0, 0), (0, 1), (2, 7)], tuning=tun)
    g = api.Graph.sbm(2000, 10, 18 / 199, 2 / 1800, seed=0)
    bf16_batch_parity(g, 64, 256, [(0, 0), (1, 7)], tuning=_t())


def test_bf16_chain_equals_single_steps_slice16():
    g = api.Graph.sbm(1500, 6, 0.02, 0.001, seed=4)
    X = api.init_uniform(g.n, 128, seed=8)
    B, T = 200, 3
    lists = api.reference_negatives(g.n, T + 7, 5)[0]
    res = []
    for mode in ("graph", "steps"):
        ctx = _ctx(g, 128, X, B, lists, gather=api.BF16, tuning=_t())
        if mode == "steps":
            for i in range(T):
                for j in range(8):
                    ctx.step(i, j)
        else:
            ctx.train(T, B, 5, 0.02)
        res.append(ctx.download_x())
        ctx.close()
    np.testing.assert_array_equal(res[0], res[1])


def test_slice_arguments():
    g = api.Graph.sbm(300, 3, 0.05, 0.01, seed=1)
    X = api.init_uniform(g.n, 8, seed=1)
    ctx = api.ForceEmbed(g.n, 8)
    try:
        with pytest.raises(api.FeError):
            ctx.set_tuning(slice=16)   # more components per lane than d = 8
        with pytest.raises(api.FeError):
            ctx.set_tuning(slice=12)
        ctx.upload_csr(g)
        ctx.upload_x(X)
        ctx.configure(64, 0.02, 5.0)
        assert ctx.get_tuning()["slice"] == 8
    finally:
        ctx.close()


GiB = 1 << 30


def _engine_specs():
    specs = [  # (graph, d, gather, batch, epochs)
        (api.Graph.rmat(12, 16, seed=1), 128, api.FP32, 512, 3),
        (api.Graph.sbm(3000, 12, 0.01, 0.0005, seed=2), 64, api.BF16, 1024, 4),
        (api.Graph.sbm(1200, 4, 0.02, 0.001, seed=3), 32, api.FP32, 256, 3),
    ]
    return [dict(g=g, x0=api.init_uniform(g.n, d, seed=20 + k), B=B, T=T, gather=gather)
            for k, (g, d, gather, B, T) in enumerate(specs)]


def test_engine_rounds_in_slices_of_16_equal_contexts():
    # one round mixes FP32 and BF16 jobs of d = 128, 64 and 32 (the run-time
    # dtype and width form) in slices of 16; each job equals the single-graph
    # context with the same tiling, through rounds and the captured schedule
    specs = _engine_specs()
    outs = {}
    for mode in ("rounds", "graph"):
        with api.Engine(2 * GiB) as e:
            e.set_tuning(slice=16)
            ids = [e.submit(s["g"], s["x0"], s["B"], s["T"], gather=s["gather"], ns=5, lr=0.02)
                   for s in specs]
            if mode == "rounds":
                while e.run(16) > 0:
                    pass
            else:
                e.run_graph()
            outs[mode] = [e.collect(i) for i in ids]
    for k, s in enumerate(specs):
        g, x0 = s["g"], s["x0"]
        with api.ForceEmbed(g.n, x0.shape[1], gather=s["gather"]) as ctx:
            ctx.upload_csr(g)
            ctx.upload_x(x0)
            ctx.set_tuning(slice=16)
            ctx.train(s["T"], s["B"], 5, 0.02)
            xc = ctx.download_x()
        assert np.isfinite(xc).all()
        np.testing.assert_array_equal(outs["rounds"][k], xc, err_msg=f"job {k}")
        np.testing.assert_array_equal(outs["graph"][k], xc, err_msg=f"job {k}")


def test_engine_slice_guards():
    g = api.Graph.sbm(600, 3, 0.03, 0.002, seed=4)
    x8 = api.init_uniform(g.n, 8, seed=1)
    x64 = api.init_uniform(g.n, 64, seed=1)
    with api.Engine(GiB) as e:
        e.set_tuning(slice=16)
        with pytest.raises(api.FeError):
            e.submit(g, x8, 128, 2, ns=5, lr=0.02)   # d below the slice
        i = e.submit(g, x64, 128, 2, ns=5, lr=0.02)
        with pytest.raises(api.FeError):
            e.set_tuning(slice=8)                   # a held job keeps its slice
        e.set_tuning(unroll=4)                      # launch shapes may change
        while e.run(16) > 0:
            pass
        x = e.collect(i)
        assert np.isfinite(x).all()
        e.set_tuning(slice=8)
        i = e.submit(g, x8, 128, 2, ns=5, lr=0.02)
        while e.run(16) > 0:
            pass
        assert np.isfinite(e.collect(i)).all()


def test_new_batch_size_drops_per_batch_tunings():
    # per-batch cuttings (mixed slices) name batches of one batch size; a new
    # batch size plans with the context's tuning, as a fresh context does
    g = api.Graph.rmat(12, 16, seed=5)
    X = api.init_uniform(g.n, 64, seed=3)
    lists = api.reference_negatives(g.n, 40, 5)[0]
    nb = -(-g.n // 512)
    ctx = _ctx(g, 64, X, 512, lists)
    ctx.set_batch_tunings([(2, 4, 1, 16, 64, 8, 16 if b % 2 else 8) for b in range(nb)])
    ctx.step_delta(0, 1, 512)
    ctx.configure(256, 0.02, 5.0)
    assert {t["slice"] for t in ctx.batch_tunings()} == {8}
    got = [ctx.step_delta(0, j, 256) for j in (0, 1, 3)]
    ctx.close()
    ref = _ctx(g, 64, X, 256, lists)
    try:
        for k, j in enumerate((0, 1, 3)):
            np.testing.assert_array_equal(got[k], ref.step_delta(0, j, 256))
    finally:
        ref.close()
