"""Writes the committed fixtures tests/golden/*.npz (small; inputs and the
oracle's outputs) with the oracle (oracle/oracle.cpp). The reference ships no
golden vectors for this path, so these pin the oracle against regressions and
give the GPU tests fixed vectors independent of the generators.

  python tests/golden/make_golden.py

Each file holds the inputs (CSR row_ptr/col_idx, X0, negative lists, sizes)
and the outputs: X after training in FP64 and FP32 (reference quirks), the
BF16-shadow form, the first batch's FP64 delta, and lr * A' X0 (SpMM) in FP64.
Loaded with numpy.load(allow_pickle=False).
"""
import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parent.parent))
import _oracle as orc  # noqa: E402
from distembed_amd import api  # noqa: E402

CASES = {
    # name: (graph, d, B, epochs, ns, lr)
    "sbm300_d16": (lambda: api.Graph.sbm(300, 6, 0.08, 0.004, seed=11), 16, 64, 3, 5, 0.02),
    "rmat8_d16": (lambda: api.Graph.rmat(8, 8, seed=12), 16, 100, 2, 3, 0.05),
}


def make(name):
    gfun, d, B, T, ns, lr = CASES[name]
    g = gfun()
    X0 = np.random.default_rng(len(name)).uniform(-0.5, 0.5, (g.n, d))
    batches = -(-g.n // B)
    lists = api.reference_negatives(g.n, T + batches - 1, ns)[0]
    x64 = X0.copy()
    orc.train(g, x64, B, T, lists, lr=lr)
    x32 = X0.astype(np.float32)
    orc.train(g, x32, B, T, lists, lr=lr)
    xbf = X0.astype(np.float32)
    orc.train(g, xbf, B, T, lists, lr=lr, bf16=True)
    delta0 = orc.step(g, X0.copy(), B, 0, 0, lists, lr=lr, apply=False, want_delta=True,
                      dtype=np.float64)
    y64 = orc.spmm(g, X0, lr)
    return dict(n=np.int64(g.n), d=np.int64(d), B=np.int64(B), epochs=np.int64(T),
                ns=np.int64(ns), lr=np.float64(lr),
                row_ptr=np.asarray(g.row_ptr, np.int64), col_idx=np.asarray(g.col_idx, np.int32),
                x0=X0, lists=np.asarray(lists, np.int32), x_fp64=x64, x_fp32=x32, x_bf16=xbf,
                delta0_fp64=np.asarray(delta0, np.float64), spmm_fp64=y64)


if __name__ == "__main__":
    for name in CASES:
        np.savez_compressed(HERE / f"{name}.npz", **make(name))
        print("wrote", HERE / f"{name}.npz")
