"""ctypes binding of libforce_embed.so (include/force_embed.h).

The library is built in-tree by distembed_amd/build.py. There is no Python or
CPU fallback for the force step: if the library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libforce_embed.so"

FE_OK = 0
FE_ERR_INVALID_ARG = 1
FE_ERR_OOM = 2
FE_ERR_HIP = 3
FE_ERR_OVER_BUDGET = 4
FE_ERR_STATE = 5
FE_ERR_NO_DEVICE = 6
FE_ERR_TIMEOUT = 7
FE_ERR_NOT_FINITE = 8

FE_GATHER_FP32 = 0
FE_GATHER_BF16 = 1
FE_QUIRK_STICKY_REP = 1
FE_QUIRK_SKIP_LAST_COL = 2
FE_QUIRKS_REFERENCE = 3
FE_NEG_STRICT = 0
FE_NEG_REDRAW = 1

_STATUS_NAMES = {
    0: "FE_OK", 1: "FE_ERR_INVALID_ARG", 2: "FE_ERR_OOM", 3: "FE_ERR_HIP",
    4: "FE_ERR_OVER_BUDGET", 5: "FE_ERR_STATE", 6: "FE_ERR_NO_DEVICE",
    7: "FE_ERR_TIMEOUT", 8: "FE_ERR_NOT_FINITE",
}


class FeError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{_STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class fe_graph(C.Structure):
    _fields_ = [("n", C.c_int64), ("nnz", C.c_int64),
                ("row_ptr", C.POINTER(C.c_int64)), ("col_idx", C.POINTER(C.c_int32))]


class fe_tuning(C.Structure):
    _fields_ = [("chunk", C.c_int), ("max_chunks", C.c_int), ("unroll", C.c_int),
                ("waves_per_block", C.c_int), ("items_per_wave", C.c_int), ("light", C.c_int),
                ("slice", C.c_int)]


class fe_job_desc(C.Structure):
    _fields_ = [("n", C.c_int64), ("d", C.c_int), ("gather_dtype", C.c_int),
                ("quirks", C.c_uint32), ("batch_size", C.c_int), ("iterations", C.c_int),
                ("ns", C.c_int), ("lr", C.c_float), ("max_bound", C.c_float),
                ("row_ptr", C.POINTER(C.c_int64)), ("col_idx", C.POINTER(C.c_int32)),
                ("nnz", C.c_int64), ("x0", C.POINTER(C.c_float)),
                ("neg_lists", C.POINTER(C.c_int32))]


_P = C.c_void_p
_I64P = C.POINTER(C.c_int64)
_I32P = C.POINTER(C.c_int32)
_FP = C.POINTER(C.c_float)
_DP = C.POINTER(C.c_double)

# name -> (restype, argtypes)
SIGNATURES = {
    "fe_create": (C.c_int, [C.POINTER(_P), C.c_int, C.c_int64, C.c_int, C.c_int, C.c_uint32]),
    "fe_destroy": (None, [_P]),
    "fe_last_error": (C.c_char_p, [_P]),
    "fe_upload_csr": (C.c_int, [_P, _I64P, _I32P, C.c_int64]),
    "fe_upload_x": (C.c_int, [_P, _FP]),
    "fe_download_x": (C.c_int, [_P, _FP]),
    "fe_set_negatives": (C.c_int, [_P, _I32P, C.c_int, C.c_int]),
    "fe_configure": (C.c_int, [_P, C.c_int, C.c_float, C.c_float]),
    "fe_set_tuning": (C.c_int, [_P, C.POINTER(fe_tuning)]),
    "fe_get_tuning": (C.c_int, [_P, C.POINTER(fe_tuning)]),
    "fe_step": (C.c_int, [_P, C.c_int, C.c_int]),
    "fe_step_delta": (C.c_int, [_P, C.c_int, C.c_int, _FP]),
    "fe_step_delta_noalloc": (C.c_int, [_P, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    "fe_train": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int]),
    "fe_synchronize": (C.c_int, [_P]),
    "fe_loss": (C.c_int, [_P, C.c_uint64, C.c_int, _DP, _DP]),
    "fe_epoch_units": (C.c_int, [_P, C.c_int, _I64P, _I64P, _I64P]),
    "fe_set_timing": (C.c_int, [_P, C.c_int]),
    "fe_kernel_times": (C.c_int, [_P, _DP, _DP, _DP, _I64P]),
    "fe_round_bf16": (C.c_int, [C.c_int, _FP, C.POINTER(C.c_uint16), C.c_int64]),
    "fe_spmm": (C.c_int, [_P, C.c_int, C.c_int, C.c_float, _FP]),
    "fe_autotune": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(fe_tuning), _DP, C.c_int,
                              C.POINTER(C.c_int), C.POINTER(fe_tuning), _DP]),
    "fe_batch_tuning": (C.c_int, [_P, C.c_int, C.POINTER(fe_tuning)]),
    "fe_set_batch_tuning": (C.c_int, [_P, C.c_int, C.POINTER(fe_tuning)]),
    "fe_reference_negatives": (C.c_int, [C.c_int64, C.c_int, C.c_int, C.c_int, _I32P, _I64P]),
    "fe_eval_negative": (C.c_int64, [C.c_uint64, C.c_int64, C.c_int, C.c_int64]),
    "fe_gen_rmat": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                              C.c_uint64, C.POINTER(fe_graph)]),
    "fe_gen_sbm": (C.c_int, [C.c_int64, C.c_int, C.c_double, C.c_double, C.c_uint64,
                             C.POINTER(fe_graph)]),
    "fe_csr_from_edges": (C.c_int, [C.c_int64, _I64P, _I64P, C.c_int64, C.POINTER(fe_graph)]),
    "fe_free_graph": (None, [C.POINTER(fe_graph)]),
    "fe_init_uniform": (C.c_int, [_FP, C.c_int64, C.c_float, C.c_float, C.c_uint64]),
    "fe_read_mtx": (C.c_int, [C.c_char_p, C.POINTER(fe_graph)]),
    "fe_write_embedding": (C.c_int, [C.c_char_p, _FP, C.c_int64, C.c_int]),
    # multi-job engine
    "fe_engine_create": (C.c_int, [C.POINTER(_P), C.c_int, C.c_int64]),
    "fe_engine_destroy": (None, [_P]),
    "fe_engine_last_error": (C.c_char_p, [_P]),
    "fe_engine_job_bytes": (C.c_int64, [_P, C.POINTER(fe_job_desc)]),
    "fe_engine_graph_timing": (C.c_int, [_P, _DP, _DP, _I64P]),
    "fe_engine_submit": (C.c_int, [_P, C.POINTER(fe_job_desc), C.POINTER(C.c_int)]),
    "fe_engine_run": (C.c_int, [_P, C.c_int, C.POINTER(C.c_int)]),
    "fe_engine_run_graph": (C.c_int, [_P]),
    "fe_engine_job_done": (C.c_int, [_P, C.c_int, C.POINTER(C.c_int)]),
    "fe_engine_collect": (C.c_int, [_P, C.c_int, _FP]),
    "fe_engine_set_tuning": (C.c_int, [_P, C.POINTER(fe_tuning)]),
    "fe_engine_stats": (C.c_int, [_P, _I64P, _I64P, _I64P]),
    "fe_engine_autotune": (C.c_int, [_P, C.POINTER(fe_job_desc), C.c_int, C.c_int,
                                     C.POINTER(fe_tuning), _DP, C.c_int, C.POINTER(C.c_int),
                                     C.POINTER(fe_tuning), _DP]),
}

_lib = None


def load(path: str | os.PathLike | None = None) -> C.CDLL:
    """Loads the native library; raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise FileNotFoundError(
            f"{p} is missing: build it with `python -m distembed_amd.build` "
            "(there is no CPU fallback for the force step)")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def device_count() -> int:
    """HIP devices visible to this process (0 without a GPU). Uses the HIP
    runtime that libforce_embed.so itself links."""
    load()
    try:
        hip = C.CDLL("libamdhip64.so.7")
    except OSError:
        return 0
    n = C.c_int(0)
    if hip.hipGetDeviceCount(C.byref(n)) != 0:
        return 0
    return n.value


def check(status: int, msg_fn) -> None:
    if status != FE_OK:
        raise FeError(status, msg_fn())
