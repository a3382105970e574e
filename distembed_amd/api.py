"""Python view of the native library (test and benchmark harness).

The product path is C++ host code plus HIP kernels behind the C ABI in
include/force_embed.h; this module only marshals numpy arrays through ctypes.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import FeError, fe_graph, fe_tuning

FP32 = _lib.FE_GATHER_FP32
BF16 = _lib.FE_GATHER_BF16
QUIRKS_REFERENCE = _lib.FE_QUIRKS_REFERENCE


def _lib_or_raise():
    return _lib.load()


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def _tuning(chunk, max_chunks, unroll, waves_per_block, items_per_wave, light, slice=None):
    """fe_tuning from optional fields: 0 keeps a field, light -1 sets 0."""
    z = lambda v: 0 if v is None else int(v)
    lt = 0 if light is None else (-1 if int(light) == 0 else int(light))
    return fe_tuning(z(chunk), z(max_chunks), z(unroll), z(waves_per_block), z(items_per_wave), lt,
                     z(slice))


def _host_err() -> str:
    return _lib_or_raise().fe_last_error(None).decode()


@dataclass
class Graph:
    """Symmetric CSR (row_ptr int64[n+1], col_idx int32[nnz])."""
    n: int
    row_ptr: np.ndarray
    col_idx: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.col_idx.shape[0])

    @staticmethod
    def _take(g: fe_graph) -> "Graph":
        lib = _lib_or_raise()
        n, nnz = int(g.n), int(g.nnz)
        rp = np.ctypeslib.as_array(g.row_ptr, shape=(n + 1,)).copy()
        ci = (np.ctypeslib.as_array(g.col_idx, shape=(nnz,)).copy() if nnz > 0
              else np.zeros(0, dtype=np.int32))
        lib.fe_free_graph(C.byref(g))
        return Graph(n, rp, ci)

    @staticmethod
    def rmat(scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19,
             c: float = 0.19, seed: int = 0) -> "Graph":
        lib = _lib_or_raise()
        g = fe_graph()
        _lib.check(lib.fe_gen_rmat(scale, edge_factor, a, b, c, seed, C.byref(g)), _host_err)
        return Graph._take(g)

    @staticmethod
    def sbm(n: int, blocks: int, p_in: float, p_out: float, seed: int = 0) -> "Graph":
        lib = _lib_or_raise()
        g = fe_graph()
        _lib.check(lib.fe_gen_sbm(n, blocks, p_in, p_out, seed, C.byref(g)), _host_err)
        return Graph._take(g)

    @staticmethod
    def from_edges(n: int, u, v) -> "Graph":
        lib = _lib_or_raise()
        u = np.ascontiguousarray(u, dtype=np.int64)
        v = np.ascontiguousarray(v, dtype=np.int64)
        g = fe_graph()
        _lib.check(lib.fe_csr_from_edges(n, _ptr(u, C.c_int64), _ptr(v, C.c_int64), u.shape[0],
                                          C.byref(g)), _host_err)
        return Graph._take(g)

    @staticmethod
    def read_mtx(path: str) -> "Graph":
        lib = _lib_or_raise()
        g = fe_graph()
        _lib.check(lib.fe_read_mtx(str(path).encode(), C.byref(g)), _host_err)
        return Graph._take(g)


def init_uniform(n: int, d: int, lo: float = -0.5, hi: float = 0.5, seed: int = 0) -> np.ndarray:
    lib = _lib_or_raise()
    x = np.empty((n, d), dtype=np.float32)
    _lib.check(lib.fe_init_uniform(_ptr(x, C.c_float), n * d, lo, hi, seed), _host_err)
    return x


def round_bf16(x: np.ndarray, device: int = 0) -> np.ndarray:
    """x rounded to BF16 on the device as the BF16 shadow holds it
    (fe_round_bf16), returned as float32."""
    lib = _lib_or_raise()
    x = np.ascontiguousarray(x, dtype=np.float32)
    bits = np.empty(x.shape, dtype=np.uint16)
    _lib.check(lib.fe_round_bf16(device, _ptr(x, C.c_float), _ptr(bits, C.c_uint16), x.size),
               _host_err)
    return (bits.astype(np.uint32) << 16).view(np.float32)


def reference_negatives(n: int, n_lists: int, ns: int, policy: int = _lib.FE_NEG_REDRAW):
    """Lists l = 0..n_lists-1 of generate_random_numbers(0, n, l, ns)
    (cpp/core/common.cpp:37-53); returns (int32 [n_lists, ns], draws equal to n)."""
    lib = _lib_or_raise()
    out = np.empty((n_lists, ns), dtype=np.int32)
    hits = C.c_int64(0)
    _lib.check(lib.fe_reference_negatives(n, n_lists, ns, policy, _ptr(out, C.c_int32),
                                          C.byref(hits)), _host_err)
    return out, int(hits.value)


def eval_negative(seed: int, r: int, k: int, n: int) -> int:
    return int(_lib_or_raise().fe_eval_negative(seed, r, k, n))


def write_embedding(path: str, x: np.ndarray) -> None:
    lib = _lib_or_raise()
    x = np.ascontiguousarray(x, dtype=np.float32)
    _lib.check(lib.fe_write_embedding(str(path).encode(), _ptr(x, C.c_float), x.shape[0],
                                      x.shape[1]), _host_err)


class ForceEmbed:
    """One graph's device-resident embedding state (fe_ctx): n rows of any
    width d in [1, 512] (held zero-padded to 8 x 2^k on the device)."""

    def __init__(self, n: int, d: int, gather: int = FP32, quirks: int = QUIRKS_REFERENCE,
                 device: int = 0):
        self._lib = _lib_or_raise()
        self._ctx = C.c_void_p()
        st = self._lib.fe_create(C.byref(self._ctx), device, n, d, gather, quirks)
        if st != _lib.FE_OK:
            raise FeError(st, self._lib.fe_last_error(None).decode())
        self.n, self.d, self.gather, self.quirks = n, d, gather, quirks

    def _err(self) -> str:
        return self._lib.fe_last_error(self._ctx).decode()

    def _check(self, st: int) -> None:
        _lib.check(st, self._err)

    def close(self) -> None:
        if self._ctx:
            self._lib.fe_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload_csr(self, g: Graph) -> None:
        rp = np.ascontiguousarray(g.row_ptr, dtype=np.int64)
        ci = np.ascontiguousarray(g.col_idx, dtype=np.int32)
        self._check(self._lib.fe_upload_csr(self._ctx, _ptr(rp, C.c_int64), _ptr(ci, C.c_int32),
                                            ci.shape[0]))

    def upload_x(self, x: np.ndarray) -> None:
        x = np.ascontiguousarray(x, dtype=np.float32)
        assert x.shape == (self.n, self.d)
        self._check(self._lib.fe_upload_x(self._ctx, _ptr(x, C.c_float)))

    def download_x(self) -> np.ndarray:
        x = np.empty((self.n, self.d), dtype=np.float32)
        self._check(self._lib.fe_download_x(self._ctx, _ptr(x, C.c_float)))
        return x

    def set_negatives(self, lists: np.ndarray) -> None:
        lists = np.ascontiguousarray(lists, dtype=np.int32)
        n_lists, ns = (lists.shape if lists.ndim == 2 else (0, 0))
        self._check(self._lib.fe_set_negatives(self._ctx, _ptr(lists, C.c_int32), n_lists, ns))

    def configure(self, batch_size: int, lr: float = 0.02, max_bound: float = 5.0) -> None:
        self._check(self._lib.fe_configure(self._ctx, batch_size, lr, max_bound))

    def set_tuning(self, chunk=None, max_chunks=None, unroll=None, waves_per_block=None,
                   items_per_wave=None, light=None, slice=None) -> None:
        """Fields left at None keep their value; light=0 turns light rows off;
        slice is the components of a row per lane (8 or 16)."""
        t = _tuning(chunk, max_chunks, unroll, waves_per_block, items_per_wave, light, slice)
        self._check(self._lib.fe_set_tuning(self._ctx, C.byref(t)))

    def get_tuning(self) -> dict:
        t = fe_tuning()
        self._check(self._lib.fe_get_tuning(self._ctx, C.byref(t)))
        return {k: getattr(t, k) for k, _ in fe_tuning._fields_}

    def step(self, epoch: int, batch: int) -> None:
        self._check(self._lib.fe_step(self._ctx, epoch, batch))

    def step_delta(self, epoch: int, batch: int, rows: int) -> np.ndarray:
        out = np.empty((rows, self.d), dtype=np.float32)
        self._check(self._lib.fe_step_delta(self._ctx, epoch, batch, _ptr(out, C.c_float)))
        return out

    def step_delta_noalloc(self, epoch: int, batch: int, delta_dev: int, stream: int = 0) -> None:
        """fe_step_delta_noalloc: the batch's delta into device memory at
        address delta_dev, enqueued on `stream` (0: the context's stream);
        returns without waiting."""
        self._check(self._lib.fe_step_delta_noalloc(self._ctx, epoch, batch, C.c_void_p(delta_dev),
                                                    C.c_void_p(stream) if stream else None))

    def train(self, iterations: int, batch_size: int, ns: int, lr: float,
              use_graph: bool = True) -> None:
        self._check(self._lib.fe_train(self._ctx, iterations, batch_size, ns, lr,
                                       1 if use_graph else 0))

    def spmm(self, iterations: int, batch_size: int, lr: float) -> np.ndarray:
        """SpMMAlgo::algo_spmm (fe_spmm): Y = lr * A' X, X unchanged."""
        y = np.empty((self.n, self.d), dtype=np.float32)
        self._check(self._lib.fe_spmm(self._ctx, iterations, batch_size, lr, _ptr(y, C.c_float)))
        return y

    def synchronize(self) -> None:
        self._check(self._lib.fe_synchronize(self._ctx))

    def loss(self, eval_seed: int = 0, eval_ns: int = 5):
        a, r = C.c_double(), C.c_double()
        self._check(self._lib.fe_loss(self._ctx, eval_seed, eval_ns, C.byref(a), C.byref(r)))
        return a.value, r.value

    def epoch_units(self, ns: int):
        e, n0, n1 = C.c_int64(), C.c_int64(), C.c_int64()
        self._check(self._lib.fe_epoch_units(self._ctx, ns, C.byref(e), C.byref(n0), C.byref(n1)))
        return e.value, n0.value, n1.value

    def autotune(self, epochs: int = 2, search_chunking: bool = False):
        """Times candidate tilings (fe_autotune), keeps the fastest; returns
        (best tuning dict, best ms/epoch, [(tuning dict, ms/epoch), ...])."""
        cap = 512
        res = (fe_tuning * cap)()
        ms = (C.c_double * cap)()
        n = C.c_int(0)
        best = fe_tuning()
        best_ms = C.c_double(0)
        self._check(self._lib.fe_autotune(self._ctx, epochs, 1 if search_chunking else 0, res, ms,
                                          cap, C.byref(n), C.byref(best), C.byref(best_ms)))
        as_dict = lambda t: {k: getattr(t, k) for k, _ in fe_tuning._fields_}
        table = [(as_dict(res[i]), ms[i]) for i in range(min(n.value, cap))]
        return as_dict(best), best_ms.value, table

    def batch_tunings(self) -> list:
        """The tuning of every batch's force launch (per-batch shapes after
        autotune)."""
        out = []
        b = 0
        while True:
            t = fe_tuning()
            if self._lib.fe_batch_tuning(self._ctx, b, C.byref(t)) != _lib.FE_OK:
                break
            out.append({k: getattr(t, k) for k, _ in fe_tuning._fields_})
            b += 1
        return out

    def set_batch_tunings(self, per_batch) -> None:
        """Per-batch tunings: one (unroll, waves_per_block, items_per_wave,
        chunk, max_chunks, light[, slice]) per batch (slice 8 when absent)."""
        for b, row in enumerate(per_batch):
            u, w, i, ch, mc, li = row[:6]
            sl = int(row[6]) if len(row) > 6 else 8
            t = fe_tuning(int(ch), int(mc), int(u), int(w), int(i), -1 if int(li) == 0 else int(li),
                          sl)
            self._check(self._lib.fe_set_batch_tuning(self._ctx, b, C.byref(t)))

    def set_timing(self, enable) -> None:
        """True or 1: an event pair around every force launch; 2: one event
        pair around each train call (divided by its force launches); False
        or 0: off."""
        mode = int(enable) if not isinstance(enable, bool) else (1 if enable else 0)
        self._check(self._lib.fe_set_timing(self._ctx, mode))

    def kernel_times(self):
        f, x, a, n = C.c_double(), C.c_double(), C.c_double(), C.c_int64()
        self._check(self._lib.fe_kernel_times(self._ctx, C.byref(f), C.byref(x), C.byref(a),
                                              C.byref(n)))
        return {"force_ms": f.value, "fixup_ms": x.value, "apply_ms": a.value,
                "launches": n.value}


class Engine:
    """Multi-job engine (fe_engine): batching and serving forms."""

    def __init__(self, hbm_budget_bytes: int, device: int = 0):
        self._lib = _lib_or_raise()
        self._e = C.c_void_p()
        st = self._lib.fe_engine_create(C.byref(self._e), device, int(hbm_budget_bytes))
        if st != _lib.FE_OK:
            raise FeError(st, self._lib.fe_engine_last_error(None).decode())
        self._keep = {}

    def _err(self) -> str:
        return self._lib.fe_engine_last_error(self._e).decode()

    def _check(self, st: int) -> None:
        _lib.check(st, self._err)

    def close(self) -> None:
        if self._e:
            self._lib.fe_engine_destroy(self._e)
            self._e = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _desc(g: Graph, x0: np.ndarray, batch_size: int, iterations: int, ns: int = 5,
              lr: float = 0.02, gather: int = FP32, quirks: int = QUIRKS_REFERENCE,
              neg_lists=None, max_bound: float = 5.0):
        rp = np.ascontiguousarray(g.row_ptr, dtype=np.int64)
        ci = np.ascontiguousarray(g.col_idx, dtype=np.int32)
        x0 = np.ascontiguousarray(x0, dtype=np.float32)
        keep = [rp, ci, x0]
        d = _lib.fe_job_desc()
        d.n, d.d, d.gather_dtype, d.quirks = g.n, x0.shape[1], gather, quirks
        d.batch_size, d.iterations, d.ns, d.lr, d.max_bound = batch_size, iterations, ns, lr, max_bound
        d.row_ptr, d.col_idx, d.nnz = _ptr(rp, C.c_int64), _ptr(ci, C.c_int32), ci.shape[0]
        d.x0 = _ptr(x0, C.c_float)
        if neg_lists is not None:
            nl = np.ascontiguousarray(neg_lists, dtype=np.int32)
            keep.append(nl)
            d.neg_lists = _ptr(nl, C.c_int32)
        return d, keep

    def job_bytes(self, g: Graph, x0: np.ndarray, batch_size: int, iterations: int, **kw) -> int:
        d, _keep = self._desc(g, x0, batch_size, iterations, **kw)
        return int(self._lib.fe_engine_job_bytes(self._e, C.byref(d)))

    def submit(self, g: Graph, x0: np.ndarray, batch_size: int, iterations: int, **kw) -> int:
        d, keep = self._desc(g, x0, batch_size, iterations, **kw)
        jid = C.c_int(-1)
        self._check(self._lib.fe_engine_submit(self._e, C.byref(d), C.byref(jid)))
        self._keep[jid.value] = (g.n, x0.shape[1])
        del keep
        return jid.value

    def run(self, rounds: int = 0) -> int:
        running = C.c_int(0)
        self._check(self._lib.fe_engine_run(self._e, rounds, C.byref(running)))
        return running.value

    def run_graph(self) -> None:
        self._check(self._lib.fe_engine_run_graph(self._e))

    def graph_timing(self):
        p, x, r = C.c_double(), C.c_double(), C.c_int64()
        self._check(self._lib.fe_engine_graph_timing(self._e, C.byref(p), C.byref(x), C.byref(r)))
        return {"prepare_ms": p.value, "exec_ms": x.value, "rounds": r.value}

    def done(self, job_id: int) -> bool:
        d = C.c_int(0)
        self._check(self._lib.fe_engine_job_done(self._e, job_id, C.byref(d)))
        return bool(d.value)

    def collect(self, job_id: int) -> np.ndarray:
        n, dim = self._keep[job_id]
        x = np.empty((n, dim), dtype=np.float32)
        self._check(self._lib.fe_engine_collect(self._e, job_id, _ptr(x, C.c_float)))
        return x

    def set_tuning(self, chunk=None, max_chunks=None, unroll=None, waves_per_block=None,
                   items_per_wave=None, light=None, slice=None) -> None:
        """As ForceEmbed.set_tuning; chunking applies to jobs submitted later
        (a new slice needs an engine that holds no job)."""
        t = _tuning(chunk, max_chunks, unroll, waves_per_block, items_per_wave, light, slice)
        self._check(self._lib.fe_engine_set_tuning(self._e, C.byref(t)))

    def autotune(self, jobs, epochs: int = 2):
        """fe_engine_autotune over `jobs`, a list of (graph, x0, batch_size,
        kwargs of submit): returns (best tuning dict, best ms, [(tuning dict,
        ms), ...]) and keeps the best as the engine's tuning."""
        descs, keep = [], []
        for (g, x0, B, kw) in jobs:
            d, k = self._desc(g, x0, B, epochs, **kw)
            descs.append(d)
            keep.append(k)
        arr = (_lib.fe_job_desc * len(descs))(*descs)
        cap = 128
        res = (fe_tuning * cap)()
        ms = (C.c_double * cap)()
        n = C.c_int(0)
        best = fe_tuning()
        best_ms = C.c_double(0)
        self._check(self._lib.fe_engine_autotune(self._e, arr, len(descs), epochs, res, ms, cap,
                                                 C.byref(n), C.byref(best), C.byref(best_ms)))
        del keep
        as_dict = lambda t: {k: getattr(t, k) for k, _ in fe_tuning._fields_}
        return as_dict(best), best_ms.value, [(as_dict(res[i]), ms[i])
                                              for i in range(min(n.value, cap))]

    def stats(self):
        b, r, u = C.c_int64(), C.c_int64(), C.c_int64()
        self._check(self._lib.fe_engine_stats(self._e, C.byref(b), C.byref(r), C.byref(u)))
        return {"bytes_in_use": b.value, "rounds_run": r.value, "units_done": u.value}
