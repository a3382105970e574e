"""Pins the oracle (oracle/oracle.cpp) -- CPU only.

The reference holds no golden vectors or known-answer tests for this path
(its tests/ only print RNG values), so the oracle is pinned by:
  * the sampler: the library's restatement of generate_random_numbers
    (cpp/core/common.cpp:37-53) against the std:: call itself, compiled by g++;
  * the per-batch update: torch autograd of the force-model loss in FP64, an
    independent route to the same update (its negative gradient is the
    unclipped update of cpp/algo/algo.hpp:573 and :641);
  * closed forms: a 2-vertex graph, zero distance, saturation of the clamp;
  * the loop structure (cpp/algo/algo.hpp:135-192): Jacobi inside a batch,
    Gauss-Seidel across batches, the sticky repulsive count, the skipped
    column n-1, list index epoch+batch -- each by an invariant that the
    corresponding mistake breaks.
"""
import math

import numpy as np
import pytest

import _oracle as orc
from distembed_amd import api


def _graph_from_pairs(n, pairs):
    u = np.array([p[0] for p in pairs], dtype=np.int64)
    v = np.array([p[1] for p in pairs], dtype=np.int64)
    return api.Graph.from_edges(n, u, v)


def _rand_graph(n, m, seed):
    rng = np.random.default_rng(seed)
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    return api.Graph.from_edges(n, u, v)


# ---- sampler -------------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 2, 3, 10, 999, 10_000, 1_048_576, 4_194_304, 2_000_000_000])
def test_sampler_restatement_matches_std(n):
    n_lists, ns = 300, 7
    # FE_NEG_STRICT writes the raw draws (including any equal to n) and then
    # reports the hits, so the buffer is compared whatever the status.
    out = np.empty((n_lists, ns), dtype=np.int32)
    lib = api._lib_or_raise()
    import ctypes as C
    hits = C.c_int64(0)
    lib.fe_reference_negatives(n, n_lists, ns, 0, out.ctypes.data_as(C.POINTER(C.c_int32)),
                               C.byref(hits))
    for l in range(n_lists):
        ref = orc.std_sampler(0, n, l, ns)
        assert np.array_equal(out[l].astype(np.int64), ref), (n, l)


def test_sampler_seed_zero_and_hits():
    # seed 0 maps the engine state to 1 ([rand.eng.lcong]); n = 1 draws from {0, 1}
    a, hits = api.reference_negatives(1, 50, 5, policy=1)
    assert a.max() == 0 and hits > 0
    raw = [orc.std_sampler(0, 1, l, 5) for l in range(50)]
    assert sum(int((r == 1).sum()) for r in raw) == hits
    with pytest.raises(api.FeError):
        api.reference_negatives(1, 50, 5, policy=0)


def test_sampler_sbm10k_lists_hit_report():
    # SBM-10K: 200 epochs x 40 batches -> 239 distinct lists of 5 over [0, 10000]
    lists, hits = api.reference_negatives(10_000, 239, 5, policy=1)
    raw = np.stack([orc.std_sampler(0, 10_000, l, 5) for l in range(239)])
    assert hits == int((raw == 10_000).sum())
    assert lists.max() < 10_000 and lists.min() >= 0
    same = raw != 10_000
    # draws before the first hit of each list are unchanged by the redraw policy
    for l in range(239):
        k = np.argmax(~same[l]) if (~same[l]).any() else 5
        assert np.array_equal(lists[l, :k], raw[l, :k])


# ---- per-batch update vs autograd --------------------------------------------------

def _autograd_delta(g, X, B, epoch, batch, lists, lr, quirks, rep):
    import torch
    n, d = X.shape
    r0, r1 = batch * B, min((batch + 1) * B, n)
    Xd = torch.tensor(X, dtype=torch.float64)
    Xb = Xd[r0:r1].clone().requires_grad_(True)
    c_lim = n - 1 if (quirks & 2) else n
    loss = torch.zeros((), dtype=torch.float64)
    for r in range(r0, r1):
        for e in range(g.row_ptr[r], g.row_ptr[r + 1]):
            c = int(g.col_idx[e])
            if c < c_lim:
                v = Xb[r - r0] - Xd[c]
                loss = loss + torch.log1p((v * v).sum())
    eps = 1e-6
    if lists is not None and lists.size:
        neg = lists[epoch + batch]
        for t in range(rep):
            for m in neg:
                v = Xb[t] - Xd[int(m)]
                s = (v * v).sum()
                loss = loss + torch.log((1 + s) / (s + eps)) / (1 - eps)
    loss.backward()
    return (-lr * Xb.grad).numpy()


@pytest.mark.parametrize("quirks", [0, 2, 3])
@pytest.mark.parametrize("epoch,batch", [(0, 0), (0, 2), (1, 1), (3, 2)])
def test_batch_update_equals_negative_gradient(quirks, epoch, batch):
    n, d, B, ns, lr = 40, 16, 16, 5, 0.02
    g = _rand_graph(n, 120, seed=7)
    rng = np.random.default_rng(3)
    X = rng.uniform(-0.5, 0.5, (n, d))
    lists = rng.integers(0, n, (epoch + batch + 1, ns)).astype(np.int32)
    # keep negatives away from the row itself so the clamp stays inactive
    batches, last = 3, n - 2 * B
    rep = B if batch < batches - 1 else last
    if (quirks & 1) and epoch > 0:
        rep = last
    delta = orc.step(g, X.copy(), B, epoch, batch, lists, lr=lr, quirks=quirks, apply=False,
                     want_delta=True, dtype=np.float64)
    ref = _autograd_delta(g, X, B, epoch, batch, lists, lr, quirks, rep)
    np.testing.assert_allclose(delta, ref, rtol=1e-9, atol=1e-13)


# ---- closed forms ---------------------------------------------------------------

def test_two_vertex_closed_form_and_skip_last_column():
    g = _graph_from_pairs(2, [(0, 1)])
    X = np.array([[0.0, 0.0], [1.0, 0.0]])
    lr = 0.1
    # s = 1, d1 = -2/(1+1) = -1; v0 = (-1, 0) -> delta0 = lr*(1, 0); v1 = (1, 0) -> (-1, 0)
    d = orc.step(g, X.copy(), 2, 0, 0, None, lr=lr, quirks=0, apply=False, want_delta=True,
                 dtype=np.float64)
    np.testing.assert_array_equal(d, [[lr, 0.0], [-lr, 0.0]])
    # with the reference's `c < n-1` test the edge 0 -> 1 is skipped
    d = orc.step(g, X.copy(), 2, 0, 0, None, lr=lr, quirks=3, apply=False, want_delta=True,
                 dtype=np.float64)
    np.testing.assert_array_equal(d, [[0.0, 0.0], [-lr, 0.0]])


def test_zero_distance_gives_zero_force():
    g = _graph_from_pairs(3, [(0, 1)])
    X = np.array([[0.25, -0.5], [0.25, -0.5], [0.7, 0.1]])
    lists = np.array([[0, 1]], dtype=np.int32)  # row 0 repels itself and its twin
    d = orc.step(g, X.copy(), 2, 0, 0, lists, lr=0.5, quirks=0, apply=False, want_delta=True,
                 dtype=np.float64)
    np.testing.assert_array_equal(d, np.zeros((2, 2)))


def test_clamp_saturation_exact():
    # s = 1e-6: d1 = 2/((2e-6)(1+1e-6)) ~ 1e6, v*d1 ~ -1e3 -> clamped to -5
    g = _graph_from_pairs(3, [(1, 2)])
    X = np.zeros((3, 4))
    X[1, 0] = 1e-3
    lists = np.array([[1]], dtype=np.int32)
    lr = 0.02
    d = orc.step(g, X.copy(), 1, 0, 0, lists, lr=lr, bound=5.0, quirks=0, apply=False,
                 want_delta=True, dtype=np.float64)
    assert d[0, 0] == -5.0 * lr and np.all(d[0, 1:] == 0)
    d32 = orc.step(g, X.astype(np.float32), 1, 0, 0, lists, lr=lr, bound=5.0, quirks=0,
                   apply=False, want_delta=True, dtype=np.float32)
    assert d32[0, 0] == np.float32(-5.0) * np.float32(lr)


# ---- loop structure -----------------------------------------------------------------

def test_jacobi_inside_batch_gauss_seidel_across():
    g = _graph_from_pairs(2, [(0, 1)])
    X0 = np.array([[0.0, 0.0], [1.0, 0.0]])
    lr = 0.1
    one = X0.copy()
    orc.train(g, one, 2, 1, None, lr=lr, quirks=0)  # one batch: both rows see pre-batch X
    np.testing.assert_array_equal(one, [[lr, 0.0], [1 - lr, 0.0]])
    two = X0.copy()
    orc.train(g, two, 1, 1, None, lr=lr, quirks=0)  # B=1: row 1 sees the moved row 0
    x0 = lr
    s = (1 - x0) ** 2
    x1 = 1 + lr * (-2 / (1 + s)) * (1 - x0)
    np.testing.assert_allclose(two, [[x0, 0.0], [x1, 0.0]], rtol=0, atol=1e-15)


def test_sticky_repulsive_count():
    n, d, B = 10, 8, 4  # batches 3, last 2
    g = _rand_graph(n, 20, seed=1)
    rng = np.random.default_rng(0)
    X = rng.uniform(-0.5, 0.5, (n, d))
    lists = rng.integers(0, n, (8, 3)).astype(np.int32)
    for quirks, reps in ((3, 2), (2, 4)):
        with_neg = orc.step(g, X.copy(), B, 1, 0, lists, quirks=quirks, apply=False,
                            want_delta=True, dtype=np.float64)
        no_neg = orc.step(g, X.copy(), B, 1, 0, None, quirks=quirks, apply=False,
                          want_delta=True, dtype=np.float64)
        diff = np.abs(with_neg - no_neg).sum(axis=1)
        assert np.all(diff[:reps] > 0) and np.all(diff[reps:] == 0), (quirks, diff)
    # epoch 0 repels all B rows of a non-final batch even with the quirk
    with_neg = orc.step(g, X.copy(), B, 0, 0, lists, quirks=3, apply=False, want_delta=True,
                        dtype=np.float64)
    no_neg = orc.step(g, X.copy(), B, 0, 0, None, quirks=3, apply=False, want_delta=True,
                      dtype=np.float64)
    assert np.all(np.abs(with_neg - no_neg).sum(axis=1) > 0)


def test_list_index_is_epoch_plus_batch():
    n, d, B = 12, 8, 4
    g = _rand_graph(n, 30, seed=2)
    rng = np.random.default_rng(5)
    X = rng.uniform(-0.5, 0.5, (n, d))
    lists = rng.integers(0, n, (10, 4)).astype(np.int32)
    a = orc.step(g, X.copy(), B, 2, 1, lists, quirks=2, apply=False, want_delta=True,
                 dtype=np.float64)
    b = orc.step(g, X.copy(), B, 0, 1, lists[2:], quirks=2, apply=False, want_delta=True,
                 dtype=np.float64)
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("quirks", [0, 3])
def test_train_equals_sequence_of_steps(quirks):
    n, d, B, T = 37, 16, 8, 3
    g = _rand_graph(n, 90, seed=4)
    rng = np.random.default_rng(9)
    X = rng.uniform(-0.5, 0.5, (n, d)).astype(np.float32)
    batches = 5
    lists, _ = api.reference_negatives(n, T + batches - 1, 5)
    a = X.copy()
    orc.train(g, a, B, T, lists, quirks=quirks)
    b = X.copy()
    for i in range(T):
        for j in range(batches):
            orc.step(g, b, B, i, j, lists, quirks=quirks, apply=True)
    np.testing.assert_array_equal(a, b)


def test_bf16_round_matches_torch_bfloat16():
    # torch's float -> bfloat16 conversion rounds to nearest even: an
    # independent implementation of the same rounding
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(3)
    base = rng.standard_normal(4096).astype(np.float32) * np.float32(3.0)
    u = base.view(np.uint32)
    ties_even = ((u & np.uint32(0xFFFE0000)) | np.uint32(0x8000)).view(np.float32)
    ties_odd = ((u & np.uint32(0xFFFE0000)) | np.uint32(0x18000)).view(np.float32)
    special = np.array([0.0, -0.0, 1e-40, -3e-39, 3.3e38, -3.3e38, 1.0, 0.5 + 2 ** -9,
                        np.inf, -np.inf], dtype=np.float32)
    x = np.concatenate([base, ties_even, ties_odd, special])
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(orc.bf16_round(x).view(np.uint32), want.view(np.uint32))
    assert np.isnan(orc.bf16_round(np.array([np.nan], dtype=np.float32))[0])


def test_bf16_train_is_fp32_forces_at_rounded_rows():
    n, d, B, T = 37, 16, 8, 3
    g = _rand_graph(n, 90, seed=4)
    rng = np.random.default_rng(9)
    X = rng.uniform(-0.5, 0.5, (n, d)).astype(np.float32)
    batches = 5
    lists, _ = api.reference_negatives(n, T + batches - 1, 5)
    # on a matrix that BF16 holds exactly, the first batch is the FP32 batch
    Xe = orc.bf16_round(X)
    a, b = Xe.copy(), Xe.copy()
    g1 = _graph_from_pairs(n, [(r, c) for r in range(B) for c in
                               g.col_idx[g.row_ptr[r]:g.row_ptr[r + 1]]])
    orc.train(g1, a, n, 1, lists, bf16=True)
    orc.train(g1, b, n, 1, lists)
    np.testing.assert_array_equal(a, b)
    # every batch re-reads the rounded matrix, including rows that the
    # previous batch just moved, and adds its forces to the float master
    a = X.copy()
    orc.train(g, a, B, T, lists, bf16=True)
    b = X.copy()
    for i in range(T):
        for j in range(batches):
            delta = orc.step(g, orc.bf16_round(b), B, i, j, lists, apply=False, want_delta=True)
            b[j * B:j * B + delta.shape[0]] += delta
    np.testing.assert_array_equal(a, b)
    assert not np.array_equal(a, orc.bf16_round(a))  # the master keeps float precision


def test_rejects_out_of_range_negative():
    g = _rand_graph(8, 10, seed=0)
    X = np.zeros((8, 8), dtype=np.float32)
    with pytest.raises(ValueError):
        orc.step(g, X, 4, 0, 0, np.array([[8]], dtype=np.int32))


def test_fp32_oracle_tracks_fp64():
    g = api.Graph.sbm(300, 3, 0.1, 0.01, seed=1)
    X = api.init_uniform(300, 32, seed=2)
    lists, _ = api.reference_negatives(300, 12, 5)
    a = X.copy()
    b = X.astype(np.float64)
    orc.train(g, a, 64, 8, lists)
    orc.train(g, b, 64, 8, lists)
    rel = np.abs(a - b).max() / np.abs(b).max()
    assert rel < 1e-5


# ---- loss -------------------------------------------------------------------------

def test_loss_closed_form():
    g = _graph_from_pairs(2, [(0, 1)])
    X = np.array([[0.0, 0.0], [1.0, 0.0]])
    a, r = orc.loss(g, X, quirks=0, seed=11, eval_ns=3)
    assert a == pytest.approx(2 * math.log(2.0), rel=1e-15)
    exp = 0.0
    for i in range(2):
        for k in range(3):
            m = orc.eval_negative(11, i, k, 2)
            s = float(((X[i] - X[m]) ** 2).sum())
            exp += math.log((1 + s) / (s + 1e-6)) / (1 - 1e-6)
    assert r == pytest.approx(exp, rel=1e-14)
    a3, _ = orc.loss(g, X, quirks=3, seed=11, eval_ns=0)
    assert a3 == pytest.approx(math.log(2.0), rel=1e-15)  # entry 0 -> 1 skipped


def test_repulsive_loss_at_a_self_pair():
    # one vertex: every evaluation pair is the row with itself, s = 0, and
    # each term is log(1/eps)/(1-eps) with eps = 1e-6, i.e. 6 ln(10)/(1-1e-6)
    g = _graph_from_pairs(1, [])
    X = np.array([[0.3, -0.2, 0.7]])
    a, r = orc.loss(g, X, quirks=0, seed=5, eval_ns=4)
    assert a == 0.0
    assert r == pytest.approx(4 * 6 * math.log(10.0) / (1 - 1e-6), rel=1e-14)


def test_repulsive_loss_gradient_is_the_repulsive_step():
    # an independent route to the repulsive loss term: its gradient by
    # central differences of the oracle's loss must equal the repulsive
    # update of the step (itself pinned against torch autograd above), with
    # lr = 1 and no clamp. A pair (r, m) moves x_r by the step of row r
    # against negative m and pulls x_m by the opposite amount.
    n, d, seed, eval_ns = 7, 4, 3, 2
    g = _graph_from_pairs(n, [])  # no edges: the loss is its repulsive part
    rng = np.random.default_rng(1)
    X = rng.uniform(-0.6, 0.6, (n, d))
    pairs = [(r, orc.eval_negative(seed, r, k, n)) for r in range(n) for k in range(eval_ns)]
    want = np.zeros_like(X)
    for (r, m) in pairs:
        lists = np.zeros((n, 1), dtype=np.int32)
        lists[r, 0] = m
        delta = orc.step(g, X.copy(), 1, 0, r, lists, lr=1.0, bound=1e30, quirks=0,
                         apply=False, want_delta=True, dtype=np.float64)[0]
        want[r] -= delta  # dL/dx_r = -(step of row r)
        want[m] += delta  # dL/dx_m = +(step of row r)
    h = 1e-6
    got = np.zeros_like(X)
    for r in range(n):
        for k in range(d):
            xp, xm = X.copy(), X.copy()
            xp[r, k] += h
            xm[r, k] -= h
            got[r, k] = (orc.loss(g, xp, quirks=0, seed=seed, eval_ns=eval_ns)[1]
                         - orc.loss(g, xm, quirks=0, seed=seed, eval_ns=eval_ns)[1]) / (2 * h)
    assert np.abs(want).max() > 0.1
    np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-6 * np.abs(want).max())


def test_eval_negative_same_in_library_and_oracle():
    for (seed, r, k, n) in [(0, 0, 0, 10), (5, 123456, 3, 1_048_576), (2**63, 7, 9, 4_194_304)]:
        assert api.eval_negative(seed, r, k, n) == orc.eval_negative(seed, r, k, n)


# ---- SpMM sibling (SpMMAlgo::algo_spmm, cpp/algo/spmm.hpp:341-412) ----------

def _masked_csr(g, quirks):
    import scipy.sparse as sp
    rp = np.asarray(g.row_ptr, dtype=np.int64)
    ci = np.asarray(g.col_idx, dtype=np.int64)
    data = np.ones(ci.size)
    if quirks & 2:
        data[ci >= g.n - 1] = 0.0  # dst_id < dst_end_index skips column n-1
    return sp.csr_matrix((data, ci, rp), shape=(g.n, g.n))


@pytest.mark.parametrize("quirks", [3, 0])
def test_spmm_matches_scipy_csr_product(quirks):
    """lr * A' X from scipy.sparse's CSR product (a library routine) against
    the oracle's serial restatement, in FP64."""
    g = api.Graph.sbm(700, 5, 0.04, 0.004, seed=3)
    X = np.random.default_rng(1).uniform(-1, 1, (g.n, 16))
    lr = 0.37
    want = lr * (_masked_csr(g, quirks) @ X)
    got = orc.spmm(g, X, lr, quirks=quirks)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14)
    # the quirk matters on this graph: vertex n-1 has neighbours
    assert (np.asarray(g.col_idx) == g.n - 1).any()


def test_spmm_closed_form_path_graph():
    """Path 0-1-2: Y0 = lr X1, Y1 = lr (X0 + X2) (lr X0 when column 2 = n-1 is
    skipped), Y2 = lr X1."""
    g = _graph_from_pairs(3, [(0, 1), (1, 2)])
    X = np.array([[1.0, 2.0], [3.0, -4.0], [0.5, 0.25]])
    lr = 0.5
    y0 = orc.spmm(g, X, lr, quirks=0)
    np.testing.assert_array_equal(y0, [[1.5, -2.0], [0.75, 1.125], [1.5, -2.0]])
    y3 = orc.spmm(g, X, lr, quirks=3)
    np.testing.assert_array_equal(y3, [[1.5, -2.0], [0.5, 1.0], [1.5, -2.0]])


def test_spmm_fp32_tracks_fp64():
    g = api.Graph.rmat(10, 8, seed=2)
    X = np.random.default_rng(4).uniform(-0.5, 0.5, (g.n, 32))
    y64 = orc.spmm(g, X, 0.02)
    y32 = orc.spmm(g, X.astype(np.float32), 0.02)
    assert y32.dtype == np.float32
    np.testing.assert_allclose(y32, y64, rtol=1e-4, atol=1e-6)


# ---- committed fixtures (tests/golden, written by make_golden.py) ----------

def _golden():
    import pathlib
    import types
    out = []
    for p in sorted((pathlib.Path(__file__).parent / "golden").glob("*.npz")):
        z = np.load(p, allow_pickle=False)
        g = types.SimpleNamespace(n=int(z["n"]), row_ptr=z["row_ptr"], col_idx=z["col_idx"])
        out.append((p.stem, z, g))
    assert out, "no fixtures under tests/golden"
    return out


def test_oracle_reproduces_committed_golden():
    """The oracle recomputes every committed fixture bit for bit: FP64 and
    FP32 training, the BF16-shadow form, batch 0's FP64 delta, and SpMM."""
    for name, z, g in _golden():
        B, T, lr = int(z["B"]), int(z["epochs"]), float(z["lr"])
        lists = z["lists"]
        x = z["x0"].copy()
        orc.train(g, x, B, T, lists, lr=lr)
        np.testing.assert_array_equal(x, z["x_fp64"], err_msg=name)
        x = z["x0"].astype(np.float32)
        orc.train(g, x, B, T, lists, lr=lr)
        np.testing.assert_array_equal(x, z["x_fp32"], err_msg=name)
        x = z["x0"].astype(np.float32)
        orc.train(g, x, B, T, lists, lr=lr, bf16=True)
        np.testing.assert_array_equal(x, z["x_bf16"], err_msg=name)
        dl = orc.step(g, z["x0"].copy(), B, 0, 0, lists, lr=lr, apply=False, want_delta=True,
                      dtype=np.float64)
        np.testing.assert_array_equal(dl, z["delta0_fp64"], err_msg=name)
        np.testing.assert_array_equal(orc.spmm(g, z["x0"], lr), z["spmm_fp64"], err_msg=name)
        # the fixtures hold real work: training moved X, SpMM is not zero
        assert np.abs(z["x_fp64"] - z["x0"]).max() > 1e-3
        assert np.abs(z["spmm_fp64"]).max() > 0
