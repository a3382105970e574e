"""Host logic of the library (no GPU): C ABI exports, generators, CSR
construction, MatrixMarket ingest, the embedding writer, argument checks."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from distembed_amd import _lib, api

ROOT = Path(__file__).resolve().parent.parent


def _header_functions():
    names = []
    for h in sorted((ROOT / "include").glob("*.h")):
        text = h.read_text()
        names += re.findall(r"^\s*(?:fe_status|void|const char \*|int64_t)\s+\**(fe_\w+)\s*\(",
                            text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol(fe_lib):
    names = _header_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(fe_lib, n)]
    assert not missing, missing
    # and the Python binding declares a signature for each of them
    assert not [n for n in names if n not in _lib.SIGNATURES]


def test_create_without_device_reports_no_device(fe_lib):
    if _lib.device_count() > 0:
        pytest.skip("a GPU is present")
    ctx = C.c_void_p()
    st = fe_lib.fe_create(C.byref(ctx), 0, 100, 128, 0, 3)
    assert st == _lib.FE_ERR_NO_DEVICE
    assert b"no HIP device" in fe_lib.fe_last_error(None)


def _check_csr(g, symmetric=True, self_loops=False):
    rp, ci = g.row_ptr, g.col_idx
    assert rp[0] == 0 and rp[-1] == g.nnz and np.all(np.diff(rp) >= 0)
    assert ci.min(initial=0) >= 0 and ci.max(initial=0) < g.n
    rows = np.repeat(np.arange(g.n), np.diff(rp))
    for r in range(0, g.n, max(1, g.n // 200)):
        seg = ci[rp[r]:rp[r + 1]]
        assert np.all(np.diff(seg) > 0)  # sorted, no duplicates
    if not self_loops:
        assert not np.any(rows == ci)
    if symmetric:
        fwd = np.sort(rows.astype(np.int64) * g.n + ci)
        bwd = np.sort(ci.astype(np.int64) * g.n + rows)
        assert np.array_equal(fwd, bwd)


def test_rmat_small_invariants_and_determinism():
    g1 = api.Graph.rmat(14, 16, seed=3)
    g2 = api.Graph.rmat(14, 16, seed=3)
    g3 = api.Graph.rmat(14, 16, seed=4)
    _check_csr(g1)
    assert np.array_equal(g1.row_ptr, g2.row_ptr) and np.array_equal(g1.col_idx, g2.col_idx)
    assert not np.array_equal(g1.col_idx[:1000], g3.col_idx[:1000])
    deg = np.diff(g1.row_ptr)
    assert deg[0] == deg.max()  # a=0.57 puts the heaviest hub at vertex 0 (natural order)


def test_rmat20_size():
    g = api.Graph.rmat(20, 16, 0.57, 0.19, 0.19, seed=0)
    assert g.n == 1 << 20
    # 16.8 M generated edges -> about 31.4 M directed entries after
    # symmetrising, merging duplicates and dropping self-loops
    assert 3.0e7 < g.nnz < 3.3e7
    deg = np.diff(g.row_ptr)
    assert deg.max() > 40_000 and (deg == 0).sum() > 300_000


def test_sbm_block_structure():
    n, blocks = 10_000, 10
    g = api.Graph.sbm(n, blocks, 18 / 999, 2 / 9000, seed=0)
    _check_csr(g)
    rows = np.repeat(np.arange(n), np.diff(g.row_ptr))
    same = (rows // 1000) == (g.col_idx // 1000)
    intra = same.sum() / n
    inter = (~same).sum() / n
    assert 17.0 < intra < 19.0 and 1.7 < inter < 2.3, (intra, inter)


def test_csr_from_edges_symmetrises_and_merges():
    g = api.Graph.from_edges(5, [0, 1, 1, 3, 2, 4], [1, 0, 1, 2, 3, 4])
    assert list(g.row_ptr) == [0, 1, 2, 3, 4, 4]
    assert list(g.col_idx) == [1, 0, 3, 2]
    with pytest.raises(api.FeError):
        api.Graph.from_edges(3, [0], [3])


def test_mtx_ingest(tmp_path):
    p = tmp_path / "sym.mtx"
    p.write_text("%%MatrixMarket matrix coordinate pattern symmetric\n% comment\n"
                 "4 4 4\n2 1\n3 1\n3 3\n4 2\n")
    g = api.Graph.read_mtx(p)
    assert g.n == 4
    assert list(g.row_ptr) == [0, 2, 4, 6, 7]
    assert list(g.col_idx) == [1, 2, 0, 3, 0, 2, 1]  # self-loop 2-2 kept
    q = tmp_path / "gen.mtx"
    q.write_text("%%MatrixMarket matrix coordinate real general\n3 3 4\n1 2 0.5\n1 2 7\n"
                 "3 1 1.0\n2 3 2\n")
    h = api.Graph.read_mtx(q)
    assert list(h.row_ptr) == [0, 1, 2, 3] and list(h.col_idx) == [1, 2, 0]
    bad = tmp_path / "bad.mtx"
    bad.write_text("%%MatrixMarket matrix coordinate pattern general\n3 4 1\n1 2\n")
    with pytest.raises(api.FeError):
        api.Graph.read_mtx(bad)


def test_mtx_bad_headers_are_errors_not_crashes(tmp_path):
    # a header announcing far more entries than follow reserves nothing huge
    big = tmp_path / "big.mtx"
    big.write_text("%%MatrixMarket matrix coordinate pattern general\n"
                   "3 3 1000000000000000\n1 2\n")
    with pytest.raises(api.FeError, match="fewer entries"):
        api.Graph.read_mtx(big)
    # vertex ids must fit the int32 column ids
    wide = tmp_path / "wide.mtx"
    wide.write_text("%%MatrixMarket matrix coordinate pattern general\n"
                    "3000000000 3000000000 1\n1 2\n")
    with pytest.raises(api.FeError):
        api.Graph.read_mtx(wide)
    for text in ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n",
                 "%%MatrixMarket matrix coordinate real skew-symmetric\n2 2 1\n2 1 1\n",
                 "%%MatrixMarket matrix coordinate pattern general\n2 2 1\n3 1\n",
                 "%%MatrixMarket matrix coordinate pattern general\n2 2\n"):
        f = tmp_path / "x.mtx"
        f.write_text(text)
        with pytest.raises(api.FeError):
            api.Graph.read_mtx(f)


def test_sbm_rejects_graphs_beyond_the_csr_limit():
    with pytest.raises(api.FeError, match="expected edge count"):
        api.Graph.sbm(200_000, 1, 1.0, 0.0, seed=0)


def test_embedding_writer_format(tmp_path):
    x = np.array([[0.5, -1.25], [1e-6, 3.0]], dtype=np.float32)
    p = tmp_path / "emb.txt"
    api.write_embedding(p, x)
    assert p.read_text() == "1 0.50000 -1.25000\n2 0.00000 3.00000\n"


def test_embedding_writer_matches_one_formatted_stream(tmp_path):
    # more rows than one group of parallel chunks (64 x 512), widths of
    # every size up to the largest float, and infinities
    rng = np.random.default_rng(7)
    n, d = 40_000, 3
    x = (rng.standard_normal((n, d)) * 10.0 ** rng.integers(-8, 8, (n, d))).astype(np.float32)
    x[5, 1] = np.finfo(np.float32).max
    x[6, 2] = -np.finfo(np.float32).max
    x[7, 0] = np.inf
    x[8, 1] = -np.inf
    x[9, 2] = -0.0
    p = tmp_path / "emb.txt"
    api.write_embedding(p, x)
    want = "".join(f"{i + 1}" + "".join(" %.5f" % float(v) for v in row) + "\n"
                   for i, row in enumerate(x))
    assert p.read_text() == want


def test_init_uniform_range_and_determinism():
    a = api.init_uniform(1000, 64, -0.5, 0.5, seed=0)
    b = api.init_uniform(1000, 64, -0.5, 0.5, seed=0)
    c = api.init_uniform(1000, 64, -0.5, 0.5, seed=1)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.min() >= -0.5 and a.max() < 0.5
    assert abs(a.mean()) < 0.01 and abs(a.std() - (1 / 12) ** 0.5) < 0.01


def test_host_argument_checks(fe_lib):
    out = np.zeros(10, dtype=np.int32)
    p = out.ctypes.data_as(C.POINTER(C.c_int32))
    assert fe_lib.fe_reference_negatives(0, 1, 1, 1, p, None) == _lib.FE_ERR_INVALID_ARG
    assert fe_lib.fe_reference_negatives(10, 1, 1, 7, p, None) == _lib.FE_ERR_INVALID_ARG
    g = _lib.fe_graph()
    assert fe_lib.fe_gen_rmat(0, 16, 0.57, 0.19, 0.19, 0, C.byref(g)) == _lib.FE_ERR_INVALID_ARG
    assert fe_lib.fe_gen_sbm(10, 20, 0.1, 0.1, 0, C.byref(g)) == _lib.FE_ERR_INVALID_ARG
    assert fe_lib.fe_read_mtx(b"/nonexistent.mtx", C.byref(g)) == _lib.FE_ERR_INVALID_ARG


def test_mtx_large_file_parsed_in_chunks(tmp_path):
    # ~1.2 M entries (tens of chunks): comment lines, CRLF line ends, tabs and
    # signs, values, and lines past the announced count (never read); the CSR
    # must equal the one built from the same pairs, duplicates merged
    rng = np.random.default_rng(7)
    n, m = 50_000, 1_200_000
    u = rng.integers(1, n + 1, m)
    v = rng.integers(1, n + 1, m)
    lines = []
    for k in range(m):
        if k % 9973 == 0:
            lines.append("% a comment line")
        sep = "\t" if k % 3 == 0 else " "
        lines.append(f"{'+' if k % 11 == 0 else ''}{u[k]}{sep}{v[k]} {k * 0.5}" + ("\r" if k % 5 == 0 else ""))
    lines += ["1 1", "not an entry"]  # past the count: ignored
    p = tmp_path / "big.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real general\n% header comment\n"
                 f"{n} {n} {m}\n" + "\n".join(lines) + "\n")
    g = api.Graph.read_mtx(p)
    key = np.unique((u - 1) * n + (v - 1))
    rows, cols = key // n, key % n
    assert g.n == n and g.nnz == key.size
    assert np.array_equal(g.row_ptr, np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=n))]))
    assert np.array_equal(g.col_idx, cols.astype(np.int32))
    # a bad line before the count is an error, wherever its chunk falls
    bad_lines = list(lines)
    bad_lines[len(bad_lines) * 2 // 3] = "12 x"
    q = tmp_path / "bad.mtx"
    q.write_text("%%MatrixMarket matrix coordinate real general\n"
                 f"{n} {n} {m}\n" + "\n".join(bad_lines) + "\n")
    with pytest.raises(api.FeError, match="bad entry line"):
        api.Graph.read_mtx(q)


def test_ctypes_structs_match_the_c_header(tmp_path):
    """The Python binding's structs have the C header's field offsets and
    sizes (fe_tuning, fe_graph, fe_job_desc), read from a program compiled
    against include/force_embed.h."""
    import subprocess
    from distembed_amd._lib import fe_graph, fe_job_desc, fe_tuning
    structs = {"fe_tuning": fe_tuning, "fe_graph": fe_graph, "fe_job_desc": fe_job_desc}
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "force_embed.h"',
             'int main(void) {']
    for name, cls in structs.items():
        lines.append(f'  printf("{name} size %zu\\n", sizeof({name}));')
        for field, _ in cls._fields_:
            lines.append(f'  printf("{name} {field} %zu\\n", offsetof({name}, {field}));')
    lines += ['  return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    root = Path(__file__).resolve().parent.parent
    subprocess.run(["gcc", "-std=c11", f"-I{root / 'include'}", str(src), "-o", str(exe)],
                   check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True,
                               text=True).stdout.splitlines():
        name, key, val = line.split()
        got[(name, key)] = int(val)
    for name, cls in structs.items():
        assert got[(name, "size")] == C.sizeof(cls), name
        for field, _ in cls._fields_:
            assert got[(name, field)] == getattr(cls, field).offset, (name, field)
