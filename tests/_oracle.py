"""ctypes wrapper of the oracle (oracle/liboracle.so) -- test infrastructure.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
to check the GPU results; the product never does.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "liboracle.so"

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not ORACLE_LIB.exists():
        subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)
    lib = C.CDLL(str(ORACLE_LIB))
    I64P, I32P = C.POINTER(C.c_int64), C.POINTER(C.c_int32)
    FP, DP = C.POINTER(C.c_float), C.POINTER(C.c_double)
    lib.orc_std_sampler.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, I64P]
    lib.orc_std_sampler.restype = None
    lib.orc_eval_negative.argtypes = [C.c_uint64, C.c_int64, C.c_int, C.c_int64]
    lib.orc_eval_negative.restype = C.c_int64
    lib.orc_step_f32.argtypes = [C.c_int64, C.c_int, I64P, I32P, FP, C.c_int, C.c_int, C.c_int,
                                 I32P, C.c_int, C.c_int, C.c_float, C.c_float, C.c_uint, FP,
                                 C.c_int, C.c_int]
    lib.orc_step_f64.argtypes = [C.c_int64, C.c_int, I64P, I32P, DP, C.c_int, C.c_int, C.c_int,
                                 I32P, C.c_int, C.c_int, C.c_double, C.c_double, C.c_uint, DP,
                                 C.c_int, C.c_int]
    lib.orc_train_f32.argtypes = [C.c_int64, C.c_int, I64P, I32P, FP, C.c_int, C.c_int, I32P,
                                  C.c_int, C.c_int, C.c_float, C.c_float, C.c_uint, C.c_int]
    lib.orc_train_f64.argtypes = [C.c_int64, C.c_int, I64P, I32P, DP, C.c_int, C.c_int, I32P,
                                  C.c_int, C.c_int, C.c_double, C.c_double, C.c_uint, C.c_int]
    lib.orc_train_bf16.argtypes = lib.orc_train_f32.argtypes
    lib.orc_spmm_f32.argtypes = [C.c_int64, C.c_int, I64P, I32P, FP, C.c_float, C.c_uint, FP, C.c_int]
    lib.orc_spmm_f32.restype = None
    lib.orc_spmm_f64.argtypes = [C.c_int64, C.c_int, I64P, I32P, DP, C.c_double, C.c_uint, DP,
                                 C.c_int]
    lib.orc_spmm_f64.restype = None
    lib.orc_bf16_round.argtypes = [FP, FP, C.c_int64]
    lib.orc_bf16_round.restype = None
    lib.orc_loss_f64.argtypes = [C.c_int64, C.c_int, I64P, I32P, DP, C.c_uint, C.c_uint64, C.c_int,
                                 DP, DP, C.c_int]
    for f in ("orc_step_f32", "orc_step_f64", "orc_train_f32", "orc_train_f64", "orc_train_bf16",
              "orc_loss_f64"):
        getattr(lib, f).restype = C.c_int
    _lib = lib
    return lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t)) if a is not None else None


def std_sampler(lo: int, hi: int, seed: int, ns: int) -> np.ndarray:
    out = np.empty(ns, dtype=np.int64)
    load().orc_std_sampler(lo, hi, seed, ns, _p(out, C.c_int64))
    return out


def eval_negative(seed, r, k, n) -> int:
    return int(load().orc_eval_negative(seed, r, k, n))


def _lists_arg(lists):
    if lists is None or lists.size == 0:
        return np.zeros((1, 1), dtype=np.int32), 0, 0
    lists = np.ascontiguousarray(lists, dtype=np.int32)
    return lists, lists.shape[0], lists.shape[1]


def step(g, X, B, epoch, batch, lists, lr=0.02, bound=5.0, quirks=3, apply=True,
         want_delta=False, threads=0, dtype=np.float32):
    """One batch on X (in place when apply). Returns the delta when asked."""
    lib = load()
    lists, n_lists, ns = _lists_arg(lists)
    rp = np.ascontiguousarray(g.row_ptr, dtype=np.int64)
    ci = np.ascontiguousarray(g.col_idx, dtype=np.int32)
    rows = min((batch + 1) * B, g.n) - batch * B
    d = X.shape[1]
    if dtype == np.float32:
        assert X.dtype == np.float32 and X.flags.c_contiguous
        delta = np.empty((rows, d), dtype=np.float32) if want_delta else None
        st = lib.orc_step_f32(g.n, d, _p(rp, C.c_int64), _p(ci, C.c_int32), _p(X, C.c_float), B,
                              epoch, batch, _p(lists, C.c_int32), n_lists, ns, lr, bound, quirks,
                              _p(delta, C.c_float), 1 if apply else 0, threads)
    else:
        assert X.dtype == np.float64 and X.flags.c_contiguous
        delta = np.empty((rows, d), dtype=np.float64) if want_delta else None
        st = lib.orc_step_f64(g.n, d, _p(rp, C.c_int64), _p(ci, C.c_int32), _p(X, C.c_double), B,
                              epoch, batch, _p(lists, C.c_int32), n_lists, ns, lr, bound, quirks,
                              _p(delta, C.c_double), 1 if apply else 0, threads)
    if st != 0:
        raise ValueError(f"oracle step failed ({st})")
    return delta


def bf16_round(X):
    """X rounded to bfloat16 (nearest even), returned as float32."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    out = np.empty_like(X)
    load().orc_bf16_round(_p(X, C.c_float), _p(out, C.c_float), X.size)
    return out


def train(g, X, B, iterations, lists, lr=0.02, bound=5.0, quirks=3, threads=0, bf16=False):
    """Training in place. bf16: the BF16-shadow form (float X only): each
    batch's forces come from X rounded to bfloat16, added to float X."""
    lib = load()
    lists, n_lists, ns = _lists_arg(lists)
    rp = np.ascontiguousarray(g.row_ptr, dtype=np.int64)
    ci = np.ascontiguousarray(g.col_idx, dtype=np.int32)
    d = X.shape[1]
    if bf16:
        assert X.dtype == np.float32
        st = lib.orc_train_bf16(g.n, d, _p(rp, C.c_int64), _p(ci, C.c_int32), _p(X, C.c_float), B,
                                iterations, _p(lists, C.c_int32), n_lists, ns, lr, bound, quirks,
                                threads)
    elif X.dtype == np.float32:
        st = lib.orc_train_f32(g.n, d, _p(rp, C.c_int64), _p(ci, C.c_int32), _p(X, C.c_float), B,
                               iterations, _p(lists, C.c_int32), n_lists, ns, lr, bound, quirks,
                               threads)
    else:
        st = lib.orc_train_f64(g.n, d, _p(rp, C.c_int64), _p(ci, C.c_int32), _p(X, C.c_double), B,
                               iterations, _p(lists, C.c_int32), n_lists, ns, lr, bound, quirks,
                               threads)
    if st != 0:
        raise ValueError(f"oracle train failed ({st})")


def spmm(g, X, lr, quirks=3, threads=0):
    """SpMMAlgo::algo_spmm at world size 1: lr * A' X in X's dtype (float32
    or float64), accumulated serially in CSR order."""
    lib = load()
    rp = np.ascontiguousarray(g.row_ptr, dtype=np.int64)
    ci = np.ascontiguousarray(g.col_idx, dtype=np.int32)
    out = np.empty_like(X)
    if X.dtype == np.float32:
        lib.orc_spmm_f32(g.n, X.shape[1], _p(rp, C.c_int64), _p(ci, C.c_int32), _p(X, C.c_float),
                         lr, quirks, _p(out, C.c_float), threads)
    else:
        lib.orc_spmm_f64(g.n, X.shape[1], _p(rp, C.c_int64), _p(ci, C.c_int32),
                         _p(X, C.c_double), lr, quirks, _p(out, C.c_double), threads)
    return out


def loss(g, X, quirks=3, seed=0, eval_ns=5, threads=0):
    lib = load()
    rp = np.ascontiguousarray(g.row_ptr, dtype=np.int64)
    ci = np.ascontiguousarray(g.col_idx, dtype=np.int32)
    X = np.ascontiguousarray(X, dtype=np.float64)
    a, r = C.c_double(), C.c_double()
    st = lib.orc_loss_f64(g.n, X.shape[1], _p(rp, C.c_int64), _p(ci, C.c_int32),
                          _p(X, C.c_double), quirks, seed, eval_ns, C.byref(a), C.byref(r),
                          threads)
    if st != 0:
        raise ValueError("oracle loss failed")
    return a.value, r.value
