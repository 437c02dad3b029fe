"""ctypes bindings for libtdfs.so (include/tdfs.h).

The library is loaded from the package directory.  There is no fallback:
if libtdfs.so is missing, importing the device entry points raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

MAX_K = 7
LIB_PATH = Path(__file__).resolve().parent / "libtdfs.so"


class Plan(C.Structure):
    _fields_ = [
        ("k", C.c_int32),
        ("labeled", C.c_int32),
        ("sym01", C.c_int32),
        ("label", C.c_int32 * MAX_K),
        ("deg", C.c_int32 * MAX_K),
        ("num_bn", C.c_int32 * MAX_K),
        ("bn", (C.c_int32 * MAX_K) * MAX_K),
        ("cond", (C.c_int32 * MAX_K) * MAX_K),
        ("shared_lvl", C.c_int32 * MAX_K),
        ("num_bn_sh", C.c_int32 * MAX_K),
        ("bn_sh", (C.c_int32 * MAX_K) * MAX_K),
        ("adj", (C.c_int32 * MAX_K) * MAX_K),
        ("npairs", C.c_int32),
        ("pairs", (C.c_int32 * 2) * 21),
        ("multiplicity", C.c_uint64),
    ]


class Tuning(C.Structure):
    _fields_ = [
        ("waves_per_block", C.c_int32),
        ("blocks_per_cu", C.c_int32),
        ("chunk_edges", C.c_int32),
        ("timeout_us", C.c_int32),
        ("split_chunk", C.c_int32),
        ("queue_log2", C.c_int32),
        ("donate_levels", C.c_int32),
        ("reuse", C.c_int32),
        ("leaf_pair_cap", C.c_int32),
        ("leaf_batch", C.c_int32),
        ("hub_min_degree", C.c_int32),
        ("hub_budget_mb", C.c_int32),
        ("degree_order", C.c_int32),
        ("time_limit_ms", C.c_int32),
        ("cursor_order", C.c_int32),
        ("waves_per_simd", C.c_int32),
        ("hoist_min", C.c_int32),
        ("idle_deadline_ms", C.c_int32),
        ("steal", C.c_int32),
        ("merge_min", C.c_int32),
        ("leaf_memo", C.c_int32),
        ("leaf_hash", C.c_int32),
        ("ext_hash", C.c_int32),
        ("tails", C.c_int32),
        ("free_order", C.c_int32),
    ]


class Result(C.Structure):
    _fields_ = [
        ("raw", C.c_uint64),
        ("count", C.c_uint64),
        ("alg_bytes", C.c_uint64),
        ("extensions", C.c_uint64),
        ("donated", C.c_uint64),
        ("kernel_ms", C.c_double),
        ("flags", C.c_uint32),
        ("stolen", C.c_uint32),
    ]


_P = C.c_void_p
_I64P = C.POINTER(C.c_int64)
_I32P = C.POINTER(C.c_int32)

_SIGS = {
    "tdfs_last_error": (C.c_char_p, []),
    "tdfs_version": (C.c_char_p, []),
    "tdfs_pattern_parse": (C.c_int, [C.c_char_p, C.c_int32, C.POINTER(Plan)]),
    "tdfs_plan_counting_order": (C.c_int, [C.POINTER(Plan), C.POINTER(Plan), _I32P]),
    "tdfs_pattern_load": (C.c_int, [C.c_char_p, C.c_int32, C.POINTER(Plan)]),
    "tdfs_pattern_reorder": (C.c_int, [C.c_char_p, _I32P, C.c_char_p, C.c_size_t]),
    "tdfs_graph_read_meta": (C.c_int, [C.c_char_p, _I64P]),
    "tdfs_graph_read": (C.c_int, [C.c_char_p, C.c_int64, C.c_int64, _I64P, _I32P, _I32P]),
    "tdfs_graph_write": (C.c_int, [C.c_char_p, C.c_int64, C.c_int64, _I64P, _I32P, _I32P]),
    "tdfs_graph_validate": (C.c_int, [C.c_int64, C.c_int64, _I64P, _I32P]),
    "tdfs_gen_er": (C.c_int, [C.c_int32, C.c_int64, C.c_uint64, _I64P, _I32P]),
    "tdfs_gen_chung_lu": (C.c_int, [C.c_int32, C.c_int64, C.c_double, C.c_double, C.c_uint64, _I64P, _I32P]),
    "tdfs_gen_labels": (C.c_int, [C.c_int32, C.c_int32, C.c_uint64, _I32P]),
    "tdfs_gen_pattern": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_uint64, C.c_char_p, C.c_size_t]),
    "tdfs_estimate_work": (C.c_int, [C.c_int32, C.c_int64, _I64P, _I32P, _I32P, C.POINTER(Plan), C.c_int32,
                                     C.c_uint64, C.c_double, C.POINTER(C.c_double)]),
    "tdfs_graph_estimate": (C.c_int, [_P, C.POINTER(Plan), C.c_int32, C.c_uint64, C.c_double,
                                      C.POINTER(C.c_double)]),
    "tdfs_tuning_default": (None, [C.POINTER(Tuning)]),
    "tdfs_ctx_create": (C.c_int, [C.c_int32, C.POINTER(_P)]),
    "tdfs_ctx_set_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "tdfs_ctx_get_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "tdfs_ctx_set_stats": (C.c_int, [_P, C.c_int32]),
    "tdfs_ctx_info": (C.c_int, [_P, _I64P]),
    "tdfs_ctx_last_profile": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "tdfs_ctx_last_counters": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "tdfs_ctx_destroy": (None, [_P]),
    "tdfs_graph_upload": (C.c_int, [_P, C.c_int32, C.c_int64, _I64P, _I32P, _I32P, C.POINTER(_P)]),
    "tdfs_graph_free": (None, [_P]),
    "tdfs_graph_info": (C.c_int, [_P, _I64P]),
    "tdfs_graph_order": (C.c_int, [_P, _I32P]),
    "tdfs_count": (C.c_int, [_P, _P, C.POINTER(Plan), C.c_int64, C.c_int64, C.POINTER(Result)]),
    "tdfs_count_noalloc": (C.c_int, [_P, _P, C.POINTER(Plan), C.c_int64, C.c_int64, C.POINTER(Result)]),
    "tdfs_count_batch": (C.c_int, [_P, _P, C.POINTER(Plan), C.c_int32, C.POINTER(Result)]),
    "tdfs_enumerate": (C.c_int, [_P, _P, C.POINTER(Plan), C.c_int64, C.c_int64, C.POINTER(C.c_int32), C.c_uint64,
                                 C.POINTER(C.c_uint64), C.POINTER(Result)]),
    "tdfs_batch_create": (C.c_int, [_P, _P, C.POINTER(Plan), C.c_int32, C.c_int32, C.POINTER(_P)]),
    "tdfs_batch_launch": (C.c_int, [_P]),
    "tdfs_batch_done": (C.c_int, [_P]),
    "tdfs_batch_wait": (C.c_int, [_P, C.POINTER(Result)]),
    "tdfs_batch_run": (C.c_int, [_P, C.POINTER(Result)]),
    "tdfs_batch_destroy": (None, [_P]),
    "tdfs_serve_create": (C.c_int, [_P, _P, C.c_int32, C.c_int32, C.c_int32, C.c_int64, C.c_double,
                                    C.POINTER(_P)]),
    "tdfs_serve_submit": (C.c_int, [_P, C.POINTER(Plan), C.POINTER(C.c_uint64)]),
    "tdfs_serve_submit_range": (C.c_int, [_P, C.POINTER(Plan), C.c_int64, C.c_int64, C.POINTER(C.c_uint64)]),
    "tdfs_serve_poll": (C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(Result), C.POINTER(C.c_double), C.c_int32,
                                  C.POINTER(C.c_int32)]),
    "tdfs_serve_status": (C.c_int, [_P, _I64P]),
    "tdfs_serve_destroy": (None, [_P]),
}

_lib = None


def lib():
    """Loads libtdfs.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m tdfs_amd.build` first")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class TdfsError(RuntimeError):
    def __init__(self, code: int, where: str):
        msg = lib().tdfs_last_error().decode(errors="replace")
        super().__init__(f"{where} failed ({code}): {msg}")
        self.code = code


def check(code: int, where: str) -> None:
    if code != 0:
        raise TdfsError(code, where)


def i64p(a: np.ndarray):
    assert a.dtype == np.int64 and a.flags.c_contiguous
    return a.ctypes.data_as(_I64P)


def i32p(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_I32P)
