"""The dist_embed driver (C++ mirror of the reference API over the C ABI) on
the GPU: its embedding.txt against the oracle run on the same graph, X0 and
negative lists."""
import json
import subprocess

import numpy as np
import pytest

import _oracle as orc
from distembed_amd import api, build

pytestmark = pytest.mark.gpu


def _run(tmp_path, *args):
    if not build.DRIVER.exists():
        pytest.fail("dist_embed driver is not built")
    out = tmp_path / "out"
    r = subprocess.run([str(build.DRIVER), *args, "-output", str(out)], capture_output=True,
                       text=True, cwd=tmp_path, timeout=300)
    assert r.returncode == 0, r.stderr
    rows = np.loadtxt(out / "embedding.txt", dtype=np.float64, ndmin=2)
    assert np.array_equal(rows[:, 0], np.arange(1, rows.shape[0] + 1))
    return rows[:, 1:], r.stdout


def _oracle_x(g, d, B, T, ns, bf16=False):
    X = api.init_uniform(g.n, d, -1.0, 1.0, seed=0)  # DenseMat's U(-1, 1), seed 0
    batches = -(-g.n // B)
    lists = api.reference_negatives(g.n, T + batches - 1, ns)[0]
    orc.train(g, X, B, T, lists, bf16=bf16)
    return X


def _sbm(n):
    nb = n / 10
    return api.Graph.sbm(n, 10, min(1.0, 18 / (nb - 1)), min(1.0, 2 / (n - nb)), seed=0)


def test_driver_sbm_matches_oracle(tmp_path):
    n, d, B, T = 900, 64, 128, 4
    got, stdout = _run(tmp_path, "-sbm", str(n), "-dim", str(d), "-batch", str(B), "-iter",
                       str(T), "-nsamples", "5", "-lr", "0.02", "-dataset", "sbm900")
    want = _oracle_x(_sbm(n), d, B, T, 5)
    # the file holds 5 decimals: half a unit of the last place, plus FP32 rounding
    np.testing.assert_allclose(got, want, rtol=0, atol=5e-6 + 1e-4 * np.abs(want).max())
    perf = (tmp_path / "perf_output").read_text()
    assert '"algo": "Embedding"' in perf and '"Total Time"' in perf and perf.rstrip().endswith(",")
    rec = json.loads(perf.rstrip().rstrip(","))
    st = rec["perf_stats"]
    assert st["Memory usage"] > 0 and st["Achieved GB/s"] > 0 and st["Edges+Negatives/s"] > 0


@pytest.mark.parametrize("d", [2, 100])
def test_driver_widths_below_and_between_powers_of_two(tmp_path, d):
    # the reference's dimension is any compile-time constant: a 2-D embedding
    # (visualisation) and d = 100 (held padded to 8 and 128 on the device)
    n, B, T = 900, 128, 3
    got, _ = _run(tmp_path, "-sbm", str(n), "-dim", str(d), "-batch", str(B), "-iter", str(T),
                  "-nsamples", "5", "-lr", "0.02")
    assert got.shape == (n, d)
    want = _oracle_x(_sbm(n), d, B, T, 5)
    np.testing.assert_allclose(got, want, rtol=0, atol=5e-6 + 1e-4 * np.abs(want).max())


def test_driver_mtx_input_equals_generated_graph(tmp_path):
    n, d, B, T = 700, 32, 100, 3
    g = _sbm(n)
    rows = np.repeat(np.arange(n), np.diff(g.row_ptr))
    keep = rows > g.col_idx  # each undirected edge once, as a symmetric file lists it
    lines = [f"{c + 1} {r + 1}" for r, c in zip(rows[keep], g.col_idx[keep])]
    mtx = tmp_path / "g.mtx"
    mtx.write_text("%%MatrixMarket matrix coordinate pattern symmetric\n"
                   f"{n} {n} {len(lines)}\n" + "\n".join(lines) + "\n")
    a, _ = _run(tmp_path, "-input", str(mtx), "-dim", str(d), "-batch", str(B), "-iter", str(T))
    b, _ = _run(tmp_path, "-sbm", str(n), "-dim", str(d), "-batch", str(B), "-iter", str(T))
    np.testing.assert_array_equal(a, b)


def test_driver_bf16_gather_matches_bf16_oracle(tmp_path):
    n, d, B, T = 800, 128, 200, 3
    got, _ = _run(tmp_path, "-sbm", str(n), "-dim", str(d), "-batch", str(B), "-iter", str(T),
                  "-gather", "bf16")
    want = _oracle_x(_sbm(n), d, B, T, 5, bf16=True)
    np.testing.assert_allclose(got, want, rtol=0, atol=5e-6 + 1e-3 * np.abs(want).max())


def test_driver_spmm_runs_one_record_per_iteration(tmp_path):
    """-spmm 1: BaselineSpMM::execute's loop (cpp/algo/baseline_spmm.hpp:51-88),
    one perf_stats entry per iteration, no embedding file."""
    if not build.DRIVER.exists():
        pytest.fail("dist_embed driver is not built")
    r = subprocess.run([str(build.DRIVER), "-sbm", "900", "-dim", "64", "-spmm", "1", "-iter",
                        "2", "-lr", "0.5"], capture_output=True, text=True, cwd=tmp_path,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    perf = (tmp_path / "perf_output").read_text()
    assert perf.count('"Local SpMM"') == 2 and perf.count('"Edges/s"') == 2
    assert '"batch_size": 900' in perf and perf.rstrip().endswith(",")
    assert not (tmp_path / "embedding.txt").exists()


def test_driver_col_major_reads_the_transposed_csr(tmp_path):
    """-col_major 1 reads the receiver's CSR, which the reference builds
    transposed (cpp/algo/algo.hpp:131-133, cpp/dist_embed.cpp:197-200); at
    world size 1 the kernel stays row-major (:184-186), so the skip of column
    N-1 applies either way. A directed ("general") graph with edges into
    N-1 tells the two CSRs apart."""
    import types
    n, d, B, T = 500, 32, 64, 3
    rng = np.random.default_rng(4)
    pairs = {(int(a), int(b)) for a, b in rng.integers(0, n, (3000, 2)) if a != b}
    pairs |= {(k, n - 1) for k in range(0, n - 1, 7)}  # edges into the last vertex
    pairs = sorted(pairs)
    u = np.array([p[0] for p in pairs])
    v = np.array([p[1] for p in pairs])
    mtx = tmp_path / "g.mtx"
    mtx.write_text("%%MatrixMarket matrix coordinate pattern general\n"
                   f"{n} {n} {len(pairs)}\n" + "\n".join(f"{a + 1} {b + 1}" for a, b in pairs)
                   + "\n")

    def csr(rows, cols):
        order = np.lexsort((cols, rows))
        rp = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=n))]).astype(np.int64)
        return types.SimpleNamespace(n=n, row_ptr=rp, col_idx=cols[order].astype(np.int32))

    native, transposed = csr(u, v), csr(v, u)
    g_file = api.Graph.read_mtx(mtx)
    assert np.array_equal(g_file.row_ptr, native.row_ptr)
    assert np.array_equal(g_file.col_idx, native.col_idx)
    batches = -(-n // B)
    lists = api.reference_negatives(n, T + batches - 1, 5)[0]
    for flag, g in (("0", native), ("1", transposed)):
        got, _ = _run(tmp_path, "-input", str(mtx), "-dim", str(d), "-batch", str(B), "-iter",
                      str(T), "-col_major", flag)
        want = api.init_uniform(n, d, -1.0, 1.0, seed=0)
        orc.train(g, want, B, T, lists, quirks=3)
        np.testing.assert_allclose(got, want, rtol=0, atol=5e-6 + 1e-4 * np.abs(want).max(),
                                   err_msg=f"-col_major {flag}")
    other = api.init_uniform(n, d, -1.0, 1.0, seed=0)
    orc.train(native, other, B, T, lists, quirks=3)
    assert np.abs(got - other).max() > 1e-3  # the two CSRs give different embeddings
