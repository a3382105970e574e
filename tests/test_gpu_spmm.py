"""GPU parity of the SpMM sibling (fe_spmm, SpMMAlgo::algo_spmm at world size
1, cpp/algo/spmm.hpp:341-412) against the oracle, through the C ABI.

Bar (FP32), elementwise on Y = lr * A' X:
    |Y_gpu - Y_32| <= 1e-5 (|Y_32| + rms(Y_32)) + 2 |Y_32 - Y_64|,
the last term admitting the FP32 oracle's serial rounding on hub rows (the
GPU sums short runs in a fixed tree). Y does not depend on the batch size or
the launch shape: each row's summation order is fixed by its degree and the
cutting, so those results must agree bit for bit.
"""
import numpy as np
import pytest

import _oracle as orc
from distembed_amd import api

pytestmark = pytest.mark.gpu

RTOL = 1e-5


def _check(y, g, X, lr, quirks=3, what=""):
    y32 = orc.spmm(g, X.astype(np.float32), lr, quirks=quirks)
    y64 = orc.spmm(g, X.astype(np.float64), lr, quirks=quirks)
    rms = float(np.sqrt(np.mean(y32.astype(np.float64) ** 2))) + 1e-30
    bound = RTOL * (np.abs(y32) + rms) + 2 * np.abs(y32.astype(np.float64) - y64)
    err = np.abs(y.astype(np.float64) - y32)
    bad = err > bound
    assert not bad.any(), (f"{what}: {bad.sum()} entries over the bound, worst "
                           f"{(err - bound).max():.3e}")


def _spmm(g, X, B, lr, gather=api.FP32, quirks=3, tuning=None, iterations=1):
    with api.ForceEmbed(g.n, X.shape[1], gather=gather, quirks=quirks) as ctx:
        ctx.upload_csr(g)
        ctx.upload_x(X)
        ctx.configure(B, lr, 5.0)
        if tuning:
            ctx.set_tuning(**tuning)
        return ctx.spmm(iterations, B, lr)


@pytest.mark.parametrize("d", [128, 64, 32, 256, 512])
def test_spmm_matches_oracle_sbm(d):
    g = api.Graph.sbm(6000, 12, 0.01, 0.0004, seed=5)
    X = api.init_uniform(g.n, d, -0.5, 0.5, seed=2)
    _check(_spmm(g, X, 1000, 0.02), g, X, 0.02, what=f"sbm d={d}")
    # rows of degree up to 32 one per lane group (columns beside the item)
    g = api.Graph.sbm(6000, 12, 0.05, 0.0004, seed=5)
    _check(_spmm(g, X, 1000, 0.02, tuning=dict(light=32)), g, X, 0.02, what=f"sbm d={d} light 32")


@pytest.mark.parametrize("quirks", [3, 0])
def test_spmm_hub_rows_and_quirk(quirks):
    """R-MAT hub rows split into chunks (partials summed in index order)."""
    g = api.Graph.rmat(14, 16, seed=3)
    X = api.init_uniform(g.n, 128, -0.5, 0.5, seed=4)
    y = _spmm(g, X, 4096, 0.37, quirks=quirks,
              tuning=dict(chunk=16, max_chunks=64, light=8))
    _check(y, g, X, 0.37, quirks=quirks, what=f"rmat14 quirks={quirks}")


def test_spmm_chunks_above_64():
    """Rows of degree up to chunk > 64 one per wave (several 64-column
    blocks), and block items of up to 4*chunk columns."""
    g = api.Graph.rmat(13, 16, seed=3)
    X = api.init_uniform(g.n, 128, -0.5, 0.5, seed=4)
    for cut in (dict(chunk=128, max_chunks=8, light=16), dict(chunk=256, max_chunks=1024)):
        _check(_spmm(g, X, 2048, 0.37, tuning=cut), g, X, 0.37, what=f"rmat13 {cut}")


def test_spmm_independent_of_batch_size_and_launch_shape():
    g = api.Graph.rmat(13, 16, seed=8)
    X = api.init_uniform(g.n, 128, -0.5, 0.5, seed=1)
    cut = dict(chunk=32, max_chunks=64, light=16)
    ref = _spmm(g, X, g.n, 0.02, tuning=dict(cut, unroll=2, waves_per_block=4, items_per_wave=1))
    _check(ref, g, X, 0.02, what="full batch")
    for B, shape in ((1000, dict(unroll=4, waves_per_block=2, items_per_wave=2)),
                     (4096, dict(unroll=1, waves_per_block=8, items_per_wave=4)),
                     (333, dict(unroll=2, waves_per_block=4, items_per_wave=1))):
        np.testing.assert_array_equal(_spmm(g, X, B, 0.02, tuning=dict(cut, **shape)), ref)
    # iterations recompute the same Y
    np.testing.assert_array_equal(_spmm(g, X, 1000, 0.02, iterations=3,
                                        tuning=dict(cut, unroll=2)), ref)


def test_spmm_bf16_gathers_use_the_rounded_rows():
    g = api.Graph.sbm(3000, 6, 0.02, 0.001, seed=9)
    X = api.init_uniform(g.n, 128, -0.5, 0.5, seed=3)
    y = _spmm(g, X, 1024, 0.02, gather=api.BF16)
    _check(y, g, orc.bf16_round(X), 0.02, what="bf16")


def test_spmm_after_training_reads_the_rotated_layout():
    """After fe_train the rows sit in rotated slots; SpMM must read row c
    wherever it lives, matching the oracle on the downloaded X."""
    g = api.Graph.sbm(5000, 10, 0.01, 0.0005, seed=6)
    X0 = api.init_uniform(g.n, 128, -0.5, 0.5, seed=5)
    with api.ForceEmbed(g.n, 128) as ctx:
        ctx.upload_csr(g)
        ctx.upload_x(X0)
        ctx.train(3, 1024, 5, 0.02)
        X = ctx.download_x()
        y = ctx.spmm(1, 1024, 0.05)
        np.testing.assert_array_equal(ctx.download_x(), X)  # X unchanged
    _check(y, g, X, 0.05, what="after training")


def test_spmm_edge_cases():
    """No edges (Y = 0), two vertices (column n-1 skipped by the quirk),
    d = 8 (one lane per row, 64 rows per wave), and isolated rows mixed with a
    heavy row."""
    g0 = api.Graph.from_edges(500, np.array([], dtype=np.int64), np.array([], dtype=np.int64))
    X0 = api.init_uniform(500, 64, seed=1)
    np.testing.assert_array_equal(_spmm(g0, X0, 128, 0.5), np.zeros_like(X0))
    g1 = api.Graph.from_edges(2, np.array([0]), np.array([1]))
    X1 = api.init_uniform(2, 16, seed=2)
    y = _spmm(g1, X1, 1, 0.5, quirks=0)  # one row per batch
    np.testing.assert_array_equal(y, 0.5 * X1[::-1])
    y = _spmm(g1, X1, 2, 0.5)  # the quirk skips column n-1 = 1
    np.testing.assert_array_equal(y, np.stack([np.zeros(16, np.float32), 0.5 * X1[0]]))
    for gather in (api.FP32, api.BF16):
        g = api.Graph.sbm(3000, 6, 0.01, 0.001, seed=4)
        X = api.init_uniform(g.n, 8, seed=3)
        Xr = X if gather == api.FP32 else orc.bf16_round(X)
        _check(_spmm(g, X, 1000, 0.02, gather=gather), g, Xr, 0.02, what=f"d=8 gather={gather}")
    star = np.arange(1, 4000)
    g = api.Graph.from_edges(8000, np.zeros_like(star), star)  # vertex 0: degree 3999
    X = api.init_uniform(g.n, 128, seed=5)
    _check(_spmm(g, X, 2048, 0.1, tuning=dict(chunk=16, max_chunks=1024)), g, X, 0.1,
           what="star")
