"""GPU parity of the HIP force step against the oracle (through the C ABI).

The per-batch bar and its reasons are stated in tests/_parity.py: normwise
1e-5 on the delta, and elementwise 1e-5 relative plus twice the FP32
oracle's own distance from the FP64 oracle.
"""
import numpy as np
import pytest

import _oracle as orc
from _parity import (RTOL, batch_parity as _batch_parity, bf16_batch_parity,
                     check_rows as _check_rows, clamp_fractions, make_ctx as _ctx,
                     rows_of as _rows)
from distembed_amd import api

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("d", [128, 64, 32])
def test_sbm_batches(d):
    g = api.Graph.sbm(2000, 10, 18 / 199, 2 / 1800, seed=0)
    # B=256: 8 batches, the last of 208 rows (sticky repulsive count active)
    _batch_parity(g, d, 256, [(0, 0), (0, 3), (0, 7), (1, 0), (1, 7), (4, 2)])
    _batch_parity(g, d, 256, [(0, 0), (1, 7)], tuning=dict(light=32))


@pytest.mark.parametrize("d", [8, 16, 256, 512])
def test_other_widths(d):
    g = api.Graph.sbm(700, 7, 0.05, 0.004, seed=3)
    _batch_parity(g, d, 128, [(0, 0), (2, 5)], ns=5 if d <= 128 else 3)
    # degrees about 15-40: rows up to degree 32 one per lane group
    g = api.Graph.sbm(700, 7, 0.25, 0.004, seed=3)
    _batch_parity(g, d, 128, [(1, 2)], ns=5 if d <= 128 else 3, tuning=dict(light=32))


def test_hub_rows_split_into_chunks():
    g = api.Graph.rmat(12, 16, seed=5)
    # small chunks force heavy rows through block partials reduced by the
    # last chunk; light rows go one per lane group unless light = 0
    for tun in ({"chunk": 4, "max_chunks": 8}, {"chunk": 32, "max_chunks": 256},
                {"chunk": 64, "max_chunks": 1}, {"chunk": 32, "light": -1},
                {"chunk": 64, "max_chunks": 64, "light": 16},
                {"chunk": 64, "max_chunks": 64, "light": 32}):
        _batch_parity(g, 128, 512, [(0, 0), (0, 1), (3, 7)], tuning=tun)


def test_chunks_above_64():
    # chunk > 64: a wave's row runs past one 64-column block (degree up to
    # chunk), block items take rows up to 4*chunk
    g = api.Graph.rmat(12, 16, seed=5)
    deg = np.diff(g.row_ptr)
    assert (deg > 64).sum() > 10 and (deg > 256).sum() > 0
    for tun in ({"chunk": 128, "max_chunks": 8, "light": 16},
                {"chunk": 256, "max_chunks": 1024, "light": 32},
                {"chunk": 200, "max_chunks": 3, "light": 0}):
        _batch_parity(g, 128, 512, [(0, 0), (0, 1), (3, 7)], tuning=tun)
    _batch_parity(g, 64, 512, [(0, 0), (2, 2)], tuning={"chunk": 256, "max_chunks": 64})
    bf16_batch_parity(g, 128, 512, [(0, 0), (1, 7)], tuning={"chunk": 128, "max_chunks": 8})


def test_block_free_and_block_batches_in_one_epoch():
    # an SBM graph (degrees far below 64) with one hub in batch 0: batch 0's
    # launch has block items, the others run the kernel without that path
    g0 = api.Graph.sbm(2000, 8, 0.02, 0.001, seed=11)
    deg0 = np.diff(g0.row_ptr)
    u = np.repeat(np.arange(g0.n), deg0)
    hub = np.full(300, 5)
    others = np.arange(700, 1000)
    g = api.Graph.from_edges(g0.n, np.concatenate([u, hub]),
                             np.concatenate([g0.col_idx, others]))
    deg = np.diff(g.row_ptr)
    B = 256
    assert deg[:B].max() > 64 and deg[B:].max() <= 64
    _batch_parity(g, 128, B, [(0, 0), (0, 1), (1, 3), (2, 7)])
    _batch_parity(g, 64, B, [(0, 0), (1, 4)], tuning={"chunk": 16, "max_chunks": 4})
    # captured epochs give the bits of stream launches and of single steps
    X = api.init_uniform(g.n, 128, seed=2)
    lists = api.reference_negatives(g.n, 3 + 7, 5)[0]
    res = []
    for mode in ("graph", "stream", "steps"):
        ctx = _ctx(g, 128, X, B, lists)
        if mode == "steps":
            for i in range(3):
                for j in range(8):
                    ctx.step(i, j)
        else:
            ctx.train(3, B, 5, 0.02, use_graph=(mode == "graph"))
        res.append(ctx.download_x())
        ctx.close()
    np.testing.assert_array_equal(res[0], res[1])
    np.testing.assert_array_equal(res[0], res[2])


def test_quirks_off_and_no_negatives():
    g = api.Graph.sbm(1000, 4, 0.02, 0.002, seed=9)
    _batch_parity(g, 64, 300, [(0, 0), (2, 3)], quirks=0)
    _batch_parity(g, 64, 300, [(0, 0), (2, 3)], ns=0)
    _batch_parity(g, 64, 300, [(1, 3)], ns=10)


def test_edges_into_last_vertex_skipped():
    n = 64
    u = np.arange(n - 1)
    v = np.full(n - 1, n - 1)
    g = api.Graph.from_edges(n, u, v)  # star around vertex n-1
    _batch_parity(g, 32, 16, [(0, 0), (0, 3)], quirks=3)
    _batch_parity(g, 32, 16, [(0, 0), (0, 3)], quirks=0)


def test_isolated_and_empty_batches():
    n = 300
    g = api.Graph.from_edges(n, np.array([0, 5]), np.array([1, 6]))  # nearly all rows empty
    _batch_parity(g, 64, 64, [(0, 0), (0, 4), (2, 2)])


# ---- scale() (cpp/algo/algo.hpp:65-72, applied at :575-579 and :642-645) -----------

def _rep_rows(n, B, epoch, j, quirks=3):
    """considering_batch_size at (epoch, batch) (cpp/algo/algo.hpp:135,167-169)."""
    batches = -(-n // B)
    last = n - B * (batches - 1)
    return last if (j == batches - 1 or (epoch > 0 and quirks & 1)) else B


def _big_components_x(n, d, seed):
    """Rows of small components plus two large ones in [-1, 1]: distances
    near 1, where the attractive |v_k d1| = 2|v_k|/(1+s) approaches its
    maximum of 1 and the repulsive one exceeds it."""
    rng = np.random.default_rng(seed)
    X = rng.uniform(-0.01, 0.01, (n, d)).astype(np.float32)
    k = rng.integers(0, d, (n, 2))
    X[np.arange(n)[:, None], k] = rng.uniform(-1, 1, (n, 2)).astype(np.float32)
    return X


@pytest.mark.parametrize("bound", [0.5, 1.0])
def test_clamp_binds_below_two(bound):
    # a bound below 2 takes the exact attractive expressions (a bound >= 2
    # can never bind: 2|v_k|/(1+s) <= 1); at 0.5 the attractive clamp binds,
    # at 1.0 only the repulsive one can
    g = api.Graph.sbm(2000, 10, 18 / 199, 2 / 1800, seed=0)
    d, B = 64, 256
    X = _big_components_x(g.n, d, seed=11)
    lists = api.reference_negatives(g.n, 20, 5)[0]
    steps = [(0, 0), (0, 3), (1, 7)]
    for (e, j) in steps:
        fa, fr = clamp_fractions(g, X, B, j, lists[e + j], _rep_rows(g.n, B, e, j), bound)
        assert fr > 1e-3, (bound, e, j, fr)
        if bound < 1.0:
            assert fa > 0.01, (bound, e, j, fa)
        else:
            assert fa == 0.0
    _batch_parity(g, d, B, steps, X=X, bound=bound, lists=lists)


def test_clamp_binding_and_free_rows_in_one_wave():
    # every third row sits in a tight cluster; each list holds two cluster
    # rows, so the repulsive clamp binds for cluster rows (s ~ 1e-5) and not
    # for the spread rows sharing their waves (med3 and single-FMA forms in
    # one launch), at the reference's bound 5
    g = api.Graph.sbm(3000, 12, 0.01, 0.0005, seed=6)
    d, B = 128, 512
    rng = np.random.default_rng(3)
    X = rng.uniform(-0.5, 0.5, (g.n, d)).astype(np.float32)
    cl = np.arange(0, g.n, 3)
    X[cl] = rng.uniform(-2e-4, 2e-4, (cl.size, d)).astype(np.float32)
    n_lists = 12
    lists = np.stack([[cl[(7 * l) % cl.size], cl[(7 * l + 3) % cl.size], 1 + 3 * l, 2 + 3 * l,
                       3 * l + 4] for l in range(n_lists)]).astype(np.int32)
    steps = [(0, 0), (0, 4), (2, 5)]
    for (e, j) in steps:
        _, fr = clamp_fractions(g, X, B, j, lists[e + j], _rep_rows(g.n, B, e, j), 5.0)
        assert 0.05 < fr < 0.95, (e, j, fr)
    _batch_parity(g, d, B, steps, X=X, bound=5.0, lists=lists)
    _batch_parity(g, d, B, steps, X=X, bound=5.0, lists=lists,
                  tuning=dict(light=0, chunk=16, max_chunks=4))


def test_non_finite_rows_follow_the_reference():
    # scale() passes a NaN through; an infinite component makes s = inf, so
    # d1 = -0 (attractive) or 0 (repulsive) and v_k * d1 is NaN where v_k is
    # infinite and +-0 elsewhere; a row of 1e20 overflows s in FP32 while
    # every component stays finite. The GPU must give the FP32 oracle's NaNs
    # and infinities at the same places and meet the bar elsewhere.
    g = api.Graph.sbm(1200, 6, 0.02, 0.001, seed=2)
    d, B = 64, 256
    X = api.init_uniform(g.n, d, seed=4)
    nan_row, inf_row, big_row = 40, 300, 700
    X[nan_row] = np.nan
    X[inf_row, 5] = np.inf
    X[big_row] = 1e20
    lists = api.reference_negatives(g.n, 12, 5)[0]
    lists[3, 1] = inf_row   # epoch 1, batch 2: every repelled row meets +inf
    lists[4, 2] = big_row   # epoch 1, batch 3: s overflows for every repelled row
    lists[lists == nan_row] = nan_row + 1
    ctx = _ctx(g, d, X, B, lists)
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
        # the cases occurred: NaN rows, a NaN component from the infinity
        assert np.isnan(ctx.step_delta(1, 2, B)).any()
    finally:
        ctx.close()


def test_tuning_does_not_change_results():
    g = api.Graph.rmat(11, 16, seed=2)
    X = api.init_uniform(g.n, 128, seed=1)
    lists = api.reference_negatives(g.n, 20, 5)[0]
    outs = []
    for tun in ({"unroll": 1, "waves_per_block": 1, "items_per_wave": 1},
                {"unroll": 4, "waves_per_block": 4, "items_per_wave": 2},
                {"unroll": 8, "waves_per_block": 8, "items_per_wave": 4},
                {"unroll": 2, "waves_per_block": 2, "items_per_wave": 16}):
        ctx = _ctx(g, 128, X, 256, lists, tuning=tun)
        ctx.train(3, 256, 5, 0.02, use_graph=False)
        outs.append(ctx.download_x())
        ctx.close()
    for o in outs[1:]:
        np.testing.assert_array_equal(o, outs[0])


def test_per_batch_tunings():
    # per-batch launch shapes give the same bits as one shape everywhere; a
    # per-batch cutting equals the run whose cutting is uniform at each batch
    # (a batch's result depends only on its own plan); autotune leaves X as
    # it found it
    g = api.Graph.rmat(12, 16, seed=5)
    X = api.init_uniform(g.n, 64, seed=3)
    B = 512
    nb = -(-g.n // B)
    lists = api.reference_negatives(g.n, 20, 5)[0]
    rng = np.random.default_rng(0)
    shapes = [(int(rng.choice([1, 2, 4, 8])), int(rng.choice([1, 2, 4, 8])),
               int(rng.choice([1, 2, 3]))) for _ in range(nb)]
    ctx = _ctx(g, 64, X, B, lists)
    ctx.set_batch_tunings([s + (64, 1024, 16) for s in shapes])
    assert [tuple(t[k] for k in ("unroll", "waves_per_block", "items_per_wave"))
            for t in ctx.batch_tunings()] == shapes
    ctx.train(3, B, 5, 0.02)
    a = ctx.download_x()
    ctx.close()
    ctx = _ctx(g, 64, X, B, lists, tuning={"chunk": 64, "max_chunks": 1024, "light": 16})
    ctx.train(3, B, 5, 0.02)
    np.testing.assert_array_equal(a, ctx.download_x())
    ctx.close()
    # cut per batch: batch 0 with chunk 16 / max_chunks 64, the rest default
    cut = [(2, 4, 1, 16, 64, 8)] + [(2, 4, 1, 64, 1024, 16)] * (nb - 1)
    ctx = _ctx(g, 64, X, B, lists)
    ctx.set_batch_tunings(cut)
    d_mixed = ctx.step_delta(0, 0, B)
    d_mixed1 = ctx.step_delta(0, 1, B)
    ctx.close()
    ctx = _ctx(g, 64, X, B, lists, tuning={"chunk": 16, "max_chunks": 64, "light": 8})
    np.testing.assert_array_equal(d_mixed, ctx.step_delta(0, 0, B))
    ctx.close()
    ctx = _ctx(g, 64, X, B, lists, tuning={"chunk": 64, "max_chunks": 1024, "light": 16})
    np.testing.assert_array_equal(d_mixed1, ctx.step_delta(0, 1, B))
    ctx.train(2, B, 5, 0.02)
    before = ctx.download_x()
    ctx.autotune(epochs=1, search_chunking=True)
    np.testing.assert_array_equal(before, ctx.download_x())
    assert len(ctx.batch_tunings()) == nb
    ctx.close()


def test_graph_replay_equals_stream_launches_and_steps():
    g = api.Graph.sbm(1500, 6, 0.02, 0.001, seed=4)
    X = api.init_uniform(g.n, 64, seed=8)
    B, T = 200, 4
    lists = api.reference_negatives(g.n, T + 7, 5)[0]
    res = []
    for mode in ("graph", "stream", "steps"):
        ctx = _ctx(g, 64, X, B, lists)
        if mode == "steps":
            for i in range(T):
                for j in range(8):
                    ctx.step(i, j)
        else:
            ctx.train(T, B, 5, 0.02, use_graph=(mode == "graph"))
        res.append(ctx.download_x())
        ctx.close()
    np.testing.assert_array_equal(res[0], res[1])
    np.testing.assert_array_equal(res[0], res[2])


def test_kernel_timing_modes():
    """Region timing (one event pair per train call) counts every force
    launch of the call and leaves the results unchanged; the per-launch
    pairs time the same launches."""
    g = api.Graph.sbm(1500, 6, 0.02, 0.001, seed=4)
    X = api.init_uniform(g.n, 64, seed=8)
    B, T = 200, 4
    lists = api.reference_negatives(g.n, T + 7, 5)[0]
    res, times = [], []
    for mode in (0, 2, 1):
        ctx = _ctx(g, 64, X, B, lists)
        ctx.set_timing(mode)
        ctx.train(T, B, 5, 0.02, use_graph=True)
        kt = ctx.kernel_times()
        res.append(ctx.download_x())
        times.append(kt)
        ctx.close()
    np.testing.assert_array_equal(res[0], res[1])
    np.testing.assert_array_equal(res[0], res[2])
    assert times[0]["launches"] == 0
    assert times[1]["launches"] == T * 8 and times[2]["launches"] == T * 8
    assert times[1]["force_ms"] > 0 and times[2]["force_ms"] > 0
    with pytest.raises(api.FeError):
        ctx = _ctx(g, 64, X, B, lists)
        try:
            ctx.set_timing(3)
        finally:
            ctx.close()


def test_training_tracks_oracle():
    g = api.Graph.sbm(1000, 10, 18 / 99, 2 / 900, seed=0)
    X = api.init_uniform(g.n, 128, seed=0)
    B, T = 256, 10
    lists = api.reference_negatives(g.n, T + 3, 5)[0]
    ctx = _ctx(g, 128, X, B, lists)
    ctx.train(T, B, 5, 0.02)
    Xg = ctx.download_x()
    ctx.close()
    Xr = X.copy()
    orc.train(g, Xr, B, T, lists)
    rel = np.abs(Xg - Xr).max() / np.abs(Xr).max()
    assert rel < 1e-4, rel


@pytest.mark.parametrize("B", [1, 7, 300, 301, 4096])
def test_batch_size_extremes(B):
    """B = 1 (one launch per row, 300 per epoch), B = n and B > n (a single
    batch of all n rows: last = n, cpp/algo/algo.hpp:80-88), two epochs each
    against the oracle."""
    n, d, T = 300, 64, 2
    g = api.Graph.sbm(n, 3, 0.05, 0.01, seed=11)
    X = api.init_uniform(n, d, seed=4)
    lists = api.reference_negatives(n, T + -(-n // B) - 1, 5)[0]
    ctx = _ctx(g, d, X, B, lists)
    ctx.train(T, B, 5, 0.02)
    Xg = ctx.download_x()
    ctx.close()
    Xr = X.copy()
    orc.train(g, Xr, B, T, lists)
    rel = np.abs(Xg - Xr).max() / np.abs(Xr).max()
    assert rel < 1e-4, rel


@pytest.mark.parametrize("gather", [api.FP32, api.BF16])
def test_graph_without_edges(gather):
    """No CSR entries at all: every row feels only its negatives."""
    n, d, B, T = 500, 32, 128, 3
    g = api.Graph.from_edges(n, np.array([], dtype=np.int64), np.array([], dtype=np.int64))
    assert g.row_ptr[-1] == 0
    X = api.init_uniform(n, d, seed=6)
    lists = api.reference_negatives(n, T + -(-n // B) - 1, 5)[0]
    ctx = _ctx(g, d, X, B, lists, gather=gather)
    ctx.train(T, B, 5, 0.02)
    Xg = ctx.download_x()
    ctx.close()
    Xr = X.copy()
    orc.train(g, Xr, B, T, lists, bf16=(gather == api.BF16))
    assert not np.array_equal(Xr, X)  # the repulsion moved the rows
    rel = np.abs(Xg - Xr).max() / np.abs(Xr).max()
    assert rel < (1e-3 if gather == api.BF16 else 1e-4), rel


def test_loss_matches_fp64_oracle():
    g = api.Graph.rmat(12, 8, seed=1)
    X = api.init_uniform(g.n, 128, seed=3)
    ctx = _ctx(g, 128, X, 1024, None)
    a, r = ctx.loss(eval_seed=7, eval_ns=5)
    ctx.close()
    ar, rr = orc.loss(g, X.astype(np.float64), seed=7, eval_ns=5)
    assert a == pytest.approx(ar, rel=1e-6)
    assert r == pytest.approx(rr, rel=1e-6)


def test_nan_in_x_reported_by_loss():
    g = api.Graph.sbm(200, 2, 0.05, 0.01, seed=1)
    X = api.init_uniform(g.n, 32, seed=1)
    X[17, 3] = np.nan
    ctx = _ctx(g, 32, X, 64, None)
    with pytest.raises(api.FeError) as e:
        ctx.loss(0, 5)
    assert e.value.status == 8
    ctx.close()


def test_bad_arguments_return_errors():
    g = api.Graph.sbm(100, 2, 0.05, 0.01, seed=1)
    ctx = api.ForceEmbed(100, 32)
    with pytest.raises(api.FeError):
        ctx.step(0, 0)  # nothing uploaded
    ctx.upload_csr(g)
    ctx.upload_x(api.init_uniform(100, 32))
    with pytest.raises(api.FeError):
        ctx.set_negatives(np.array([[100]], dtype=np.int32))
    with pytest.raises(api.FeError):
        ctx.configure(0)
    ctx.configure(32)
    ctx.set_negatives(np.zeros((2, 3), dtype=np.int32))
    with pytest.raises(api.FeError):
        ctx.step(5, 1)  # list 6 does not exist
    with pytest.raises(api.FeError):
        ctx.step(0, 4)  # batch out of range
    for bad in (dict(light=33), dict(chunk=1025), dict(slice=4)):
        with pytest.raises(api.FeError):
            ctx.set_tuning(**bad)
    ctx.set_tuning(light=32, chunk=1024)
    ctx.close()
    for d in (0, 513):
        with pytest.raises(api.FeError):
            api.ForceEmbed(100, d)


@pytest.mark.parametrize("d", [2, 3, 12, 100, 200])
def test_any_width_against_oracle(d):
    # rows held zero-padded to 8 x 2^k on the device; the caller sees width d
    g = api.Graph.sbm(700, 7, 0.05, 0.004, seed=3)
    _batch_parity(g, d, 128, [(0, 0), (2, 5)])
    g = api.Graph.rmat(11, 16, seed=2)
    _batch_parity(g, d, 256, [(0, 0), (1, 3)], tuning={"chunk": 16, "max_chunks": 8})
    bf16_batch_parity(g, d, 256, [(0, 0), (1, 5)])


@pytest.mark.parametrize("d,gather", [(12, api.FP32), (3, api.FP32), (100, api.BF16)])
def test_padded_width_equals_zero_columns(d, gather):
    # a run at width d gives the bits of the run at the padded width whose
    # extra columns are zero, and those columns stay zero
    g = api.Graph.rmat(11, 16, seed=4)
    D = 8
    while D < d:
        D *= 2
    X = api.init_uniform(g.n, d, seed=5)
    XD = np.zeros((g.n, D), dtype=np.float32)
    XD[:, :d] = X
    lists = api.reference_negatives(g.n, 4 + 8, 5)[0]
    out = []
    for x, w in ((X, d), (XD, D)):
        ctx = api.ForceEmbed(g.n, w, gather=gather)
        ctx.upload_csr(g)
        ctx.upload_x(x)
        ctx.set_negatives(lists)
        ctx.configure(256, 0.02, 5.0)
        ctx.train(4, 256, 5, 0.02)
        out.append(ctx.download_x())
        if w == d:
            y = ctx.spmm(1, 256, 0.3)
        else:
            yD = ctx.spmm(1, 256, 0.3)
        ctx.close()
    np.testing.assert_array_equal(out[0], out[1][:, :d])
    assert not out[1][:, d:].any()
    np.testing.assert_array_equal(y, yD[:, :d])
    assert not yD[:, d:].any()



# ---- full-size workload (R-MAT scale 20, d = 128, B = 16384) ----------------------

@pytest.fixture(scope="module")
def rmat20():
    return api.Graph.rmat(20, 16, 0.57, 0.19, 0.19, seed=0)


def test_rmat20_batches_match_oracle(rmat20):
    g = rmat20
    # batch 0 holds the hub rows (max degree ~6.5e4), batch 63 the tail
    _batch_parity(g, 128, 16384, [(0, 0), (0, 1), (0, 63), (1, 17)])


def test_rmat20_epoch_properties(rmat20):
    g = rmat20
    d, B = 128, 16384
    X = api.init_uniform(g.n, d, seed=0)
    lists = api.reference_negatives(g.n, 2 + 63, 5)[0]
    ctx = _ctx(g, d, X, B, lists)
    e, n0, n1 = ctx.epoch_units(5)
    assert e == int((g.col_idx < g.n - 1).sum())
    assert n0 == n1 == g.n * 5
    ctx.train(2, B, 5, 0.02, use_graph=True)
    X1 = ctx.download_x()
    ctx.upload_x(X)
    ctx.train(2, B, 5, 0.02, use_graph=True)
    X2 = ctx.download_x()
    ctx.close()
    np.testing.assert_array_equal(X1, X2)  # bitwise reproducible
    assert np.isfinite(X1).all()
    # every row moves by at most 2 epochs x (deg + 5) x lr x bound per component
    deg = np.diff(g.row_ptr)
    step_max = 2 * (deg + 5) * 0.02 * 5.0
    assert np.all(np.abs(X1 - X).max(axis=1) <= step_max + 1e-6)


# ---- BF16 gather shadow -----------------------------------------------------------

def _train_x(g, d, X, B, T, lists, gather, use_graph=True, steps=False):
    ctx = _ctx(g, d, X, B, lists, gather=gather)
    try:
        if steps:
            batches = -(-g.n // B)
            for i in range(T):
                for j in range(batches):
                    ctx.step(i, j)
        else:
            ctx.train(T, B, lists.shape[1], 0.02, use_graph=use_graph)
        return ctx.download_x(), ctx.loss(0, 5)
    finally:
        ctx.close()


def test_bf16_chain_equals_single_steps():
    # the fused chain reads the previous batch's rows from staging rounded to
    # BF16, exactly as the shadow refresh of a separate apply would
    g = api.Graph.sbm(1500, 6, 0.02, 0.001, seed=4)
    X = api.init_uniform(g.n, 128, seed=8)
    lists = api.reference_negatives(g.n, 12, 5)[0]
    a, _ = _train_x(g, 128, X, 200, 4, lists, api.BF16)
    b, _ = _train_x(g, 128, X, 200, 4, lists, api.BF16, use_graph=False)
    c, _ = _train_x(g, 128, X, 200, 4, lists, api.BF16, steps=True)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(a, c)


def test_bf16_row_meeting_itself_feels_no_force():
    # a row drawn as its own negative: zero repulsion in the reference (v = 0);
    # with BF16 gathers the distance must use the row as the shadow holds it
    n, d = 64, 128
    g = api.Graph.from_edges(n, np.array([3]), np.array([3]))  # only a self-loop (dropped)
    X = api.init_uniform(n, d, seed=5)
    lists = np.array([[7, 7, 7]], dtype=np.int32)
    ctx = _ctx(g, d, X, 16, lists, gather=api.BF16)
    delta = ctx.step_delta(0, 0, 16)
    ctx.close()
    assert np.all(delta[7] == 0)
    assert np.all(np.abs(delta).max(axis=1)[[0, 1, 2, 4, 5, 6, 8]] > 0)


def test_bf16_gradient_error_vs_fp64():
    g = api.Graph.rmat(14, 16, seed=3)
    d, B = 128, 2048
    X = api.init_uniform(g.n, d, seed=2)
    lists = api.reference_negatives(g.n, 12, 5)[0]
    ctx = _ctx(g, d, X, B, lists, gather=api.BF16)
    try:
        for (epoch, j) in ((0, 0), (0, 3), (2, 7)):
            rows = _rows(g.n, B, j)
            dg = ctx.step_delta(epoch, j, rows).astype(np.float64)
            d64 = orc.step(g, X.astype(np.float64), B, epoch, j, lists, apply=False,
                           want_delta=True, dtype=np.float64)
            err = np.linalg.norm(dg - d64) / np.linalg.norm(d64)
            # BF16 keeps 8 significant bits: 2^-9 relative per gathered value
            assert err < 1e-2, (epoch, j, err)
    finally:
        ctx.close()


def test_bf16_final_loss_within_one_percent_of_fp32():
    # SBM (the reference's CPU workload shape): well conditioned, FP32 and FP64
    # runs agree to ~1e-8 after 20 epochs, so the BF16 gap is the BF16 effect
    g = api.Graph.sbm(10_000, 10, 18 / 999, 2 / 9000, seed=0)
    d, B, T = 128, 256, 20
    X = api.init_uniform(g.n, d, seed=0)
    lists = api.reference_negatives(g.n, T + 40, 5)[0]
    _, (a32, r32) = _train_x(g, d, X, B, T, lists, api.FP32)
    _, (a16, r16) = _train_x(g, d, X, B, T, lists, api.BF16)
    l32, l16 = a32 + r32, a16 + r16
    assert abs(l16 - l32) <= 0.01 * abs(l32), (l16, l32)


def test_device_bf16_rounding_matches_oracle():
    """The shadow's rounding on the device (fe_round_bf16, the conversion
    every BF16 launch applies) equals the oracle's round-to-nearest-even bit
    for bit: random bit patterns, exact ties with an even and an odd kept
    bit, neighbours of ties, subnormals, overflow to infinity, signed zeros
    and infinities. A NaN stays a NaN."""
    rng = np.random.default_rng(5)
    hi = rng.integers(0, 2 ** 16, 4096, dtype=np.uint64).astype(np.uint32)
    special = np.array([0x00000000, 0x80000000, 0x00000001, 0x00008000, 0x00018000,
                        0x00017FFF, 0x007FFFFF, 0x00800000, 0x807FFFFF, 0x3F808000,
                        0x3F818000, 0x7F7F7FFF, 0x7F7F8000, 0x7F7FFFFF, 0xFF7FFFFF,
                        0x7F800000, 0xFF800000, 0x7FC00000, 0x7F800001, 0xFFFFFFFF],
                       dtype=np.uint32)
    bits = np.concatenate([
        rng.integers(0, 2 ** 32, 200_000, dtype=np.uint64).astype(np.uint32),
        (hi << 16) | 0x8000, (hi << 16) | 0x7FFF, (hi << 16) | 0x8001, special])
    x = bits.view(np.float32)
    nan = np.isnan(x)
    assert nan.sum() > 100
    for m in (x.size, 7):  # an odd count too
        got = api.round_bf16(x[:m])
        want = orc.bf16_round(x[:m])
        assert np.array_equal(np.isnan(got), nan[:m])
        ok = ~nan[:m]
        assert np.array_equal(got[ok].view(np.uint32), want[ok].view(np.uint32))


def test_bf16_batches_match_oracle_at_rounded_rows():
    # the BF16 form's forces are the FP32 forces at X rounded to bfloat16
    # (every gather and the row's own distance), so on identical X they meet
    # the FP32 per-batch bound against the oracle run on the rounded matrix
    g = api.Graph.rmat(14, 16, seed=3)
    bf16_batch_parity(g, 128, 2048, [(0, 0), (0, 3), (2, 7)])
    bf16_batch_parity(g, 128, 2048, [(1, 0)],
                       tuning=dict(chunk=32, max_chunks=64, light=0, unroll=1))
    g = api.Graph.sbm(3000, 6, 0.01, 0.001, seed=5)
    bf16_batch_parity(g, 64, 512, [(0, 0), (3, 5)])
    # widths without a compile-time lane count (G = 32, 64 lanes per row)
    g = api.Graph.sbm(700, 7, 0.05, 0.004, seed=3)
    bf16_batch_parity(g, 256, 128, [(0, 0), (2, 5)])
    bf16_batch_parity(g, 512, 128, [(1, 3)])


def test_bf16_training_tracks_bf16_oracle():
    # 10 epochs of the captured chain (shadow refreshed from each batch's new
    # rows) against the oracle's BF16 form. Rows differ by FP32 rounding, which
    # now and then moves a value across a BF16 rounding boundary: a 2^-9
    # relative change of one gathered component, hence the looser max bound
    g = api.Graph.sbm(1000, 10, 18 / 99, 2 / 900, seed=0)
    X = api.init_uniform(g.n, 128, seed=0)
    B, T = 256, 10
    lists = api.reference_negatives(g.n, T + 3, 5)[0]
    ctx = _ctx(g, 128, X, B, lists, gather=api.BF16)
    ctx.train(T, B, 5, 0.02)
    Xg = ctx.download_x()
    ctx.close()
    Xr = X.copy()
    orc.train(g, Xr, B, T, lists, bf16=True)
    assert np.linalg.norm(Xg - Xr) <= 1e-5 * np.linalg.norm(Xr)
    assert np.abs(Xg - Xr).max() <= 1e-3 * np.abs(Xr).max()


def test_gpu_matches_committed_golden():
    """The HIP path against the committed fixtures (tests/golden): FP32
    training to the FP32 bar with the fixture's FP64 result as the rounding
    reference, the BF16 form as in the BF16 training test, and SpMM."""
    import pathlib
    import types
    files = sorted((pathlib.Path(__file__).parent / "golden").glob("*.npz"))
    assert files
    for p in files:
        z = np.load(p, allow_pickle=False)
        g = types.SimpleNamespace(n=int(z["n"]), row_ptr=z["row_ptr"], col_idx=z["col_idx"])
        gg = api.Graph.from_edges(g.n, np.repeat(np.arange(g.n), np.diff(g.row_ptr)), g.col_idx)
        assert np.array_equal(gg.row_ptr, g.row_ptr) and np.array_equal(gg.col_idx, g.col_idx)
        d, B, T, ns, lr = int(z["d"]), int(z["B"]), int(z["epochs"]), int(z["ns"]), float(z["lr"])
        x0 = z["x0"].astype(np.float32)
        for gather, want in ((api.FP32, z["x_fp32"]), (api.BF16, z["x_bf16"])):
            ctx = _ctx(gg, d, x0, B, z["lists"], gather=gather, lr=lr)
            if gather == api.FP32:
                y = ctx.spmm(1, B, lr)  # on X0, which SpMM leaves unchanged
            ctx.train(T, B, ns, lr)
            xg = ctx.download_x()
            ctx.close()
            if gather == api.FP32:
                bound = RTOL * (np.abs(want) + np.sqrt(np.mean(want.astype(np.float64) ** 2))) \
                    + 2 * np.abs(want.astype(np.float64) - z["x_fp64"]) + 1e-7
                assert (np.abs(xg.astype(np.float64) - want) <= bound).all(), p.stem
                y64 = z["spmm_fp64"]
                assert np.allclose(y, y64, rtol=1e-5, atol=1e-6 * np.abs(y64).max()), p.stem
            else:
                assert np.linalg.norm(xg - want) <= 1e-5 * np.linalg.norm(want), p.stem
                assert np.abs(xg - want).max() <= 1e-3 * np.abs(want).max(), p.stem
