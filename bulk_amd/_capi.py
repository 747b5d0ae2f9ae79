"""ctypes binding of the C-ABI declared in include/bulkx.h.

The library is loaded from bulk_amd/lib/libbulkx.so (built in-tree by
``bulk_amd.build``).  There is no fallback: if the library is missing the
import of any entry point raises ``BulkLibraryMissing``.
"""
from __future__ import annotations

import ctypes
import os
import re

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libbulkx.so")
HEADER_PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                           "include", "bulkx.h")

# enum values (mirrors include/bulkx.h)
INT32, UINT32, INT64, UINT64, FLOAT32, FLOAT64 = range(6)
PLUS, MIN, MAX = range(3)
LESS, GREATER = range(2)
EQUAL_TO = 0
(ALGO_INCLUSIVE_SCAN, ALGO_EXCLUSIVE_SCAN, ALGO_REDUCE, ALGO_MERGE, ALGO_MERGE_BY_KEY,
 ALGO_SORT_BY_KEY, ALGO_REDUCE_BY_KEY, ALGO_REDUCE_INTERVALS) = range(8)
# tuning-table keys of the batch executor's op families
ALGO_BATCH_SCAN, ALGO_BATCH_REDUCE_BY_KEY, ALGO_BATCH_MERGE = 8, 9, 10
(JOB_INCLUSIVE_SCAN, JOB_REDUCE_BY_KEY, JOB_MERGE_BY_KEY, JOB_SORT_BY_KEY, JOB_REDUCE) = range(5)

STATUS_NAMES = {
    0: "BULKX_SUCCESS", 1: "BULKX_ERR_INVALID_ARGUMENT", 2: "BULKX_ERR_UNSUPPORTED_DTYPE",
    3: "BULKX_ERR_UNSUPPORTED_OP", 4: "BULKX_ERR_WORKSPACE_TOO_SMALL", 5: "BULKX_ERR_NO_DEVICE",
    6: "BULKX_ERR_HIP", 7: "BULKX_ERR_OUT_OF_MEMORY", 8: "BULKX_ERR_BUDGET",
    9: "BULKX_ERR_NOT_READY", 10: "BULKX_ERR_INTERNAL", 11: "BULKX_ERR_LOOKBACK",
}
ERR_LOOKBACK = 11


class BulkLibraryMissing(RuntimeError):
    pass


class BulkError(RuntimeError):
    """A non-zero status from the C-ABI (the counterpart of the reference's
    thrust::system_error thrown by throw_on_error, bulk/detail/throw_on_error.hpp:33-49)."""

    def __init__(self, status: int, where: str, detail: str = ""):
        self.status = status
        name = STATUS_NAMES.get(status, f"status {status}")
        super().__init__(f"{where}: {name}" + (f" ({detail})" if detail else ""))


class Job(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32), ("dtype", ctypes.c_int32), ("value_dtype", ctypes.c_int32),
        ("reserved", ctypes.c_int32), ("n0", ctypes.c_int64), ("n1", ctypes.c_int64),
        ("in0", ctypes.c_void_p), ("in1", ctypes.c_void_p), ("in2", ctypes.c_void_p),
        ("in3", ctypes.c_void_p), ("out0", ctypes.c_void_p), ("out1", ctypes.c_void_p),
        ("out2", ctypes.c_void_p),
    ]


_P = ctypes.c_void_p
_I = ctypes.c_int
_L = ctypes.c_int64
_S = ctypes.c_size_t

# name -> (restype, argtypes)
SIGNATURES = {
    "bulkx_status_string": (ctypes.c_char_p, [_I]),
    "bulkx_last_error": (ctypes.c_char_p, []),
    "bulkx_version": (ctypes.c_char_p, []),
    "bulkx_dtype_size": (_S, [_I]),
    "bulkx_stream_status": (_I, [_P]),
    "bulkx_set_lookback_spin_limit": (_I, [ctypes.c_uint32]),
    "bulkx_device_info": (_I, [ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I),
                               ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_longlong)]),
    "bulkx_choose_sizes": (_I, [_I, _I, _L, ctypes.POINTER(_L), ctypes.POINTER(_I),
                                ctypes.POINTER(_I)]),
    "bulkx_load_tuning": (_I, [ctypes.c_char_p, ctypes.POINTER(_I)]),
    "bulkx_tuning_variants": (_I, [_I, ctypes.POINTER(_I)]),
    "bulkx_set_tuning": (_I, [_I, _I, _I, _I]),
    "bulkx_tuned_variant": (_I, [_I, _I, _L, ctypes.POINTER(_I)]),
    "bulkx_workspace_size": (_I, [_I, _I, _I, _L, _L, ctypes.POINTER(_S)]),
    "bulkx_inclusive_scan": (_I, [_I, _I, _P, _P, _L, _P, _P, _S, _P]),
    "bulkx_exclusive_scan": (_I, [_I, _I, _P, _P, _L, _P, _P, _S, _P]),
    "bulkx_inclusive_scan_noalloc": (_I, [_I, _I, _P, _P, _L, _P, _P]),
    "bulkx_reduce": (_I, [_I, _I, _P, _L, _P, _P, _P, _S, _P]),
    "bulkx_merge": (_I, [_I, _I, _P, _L, _P, _L, _P, _P, _S, _P]),
    "bulkx_merge_by_key": (_I, [_I, _I, _I, _P, _P, _L, _P, _P, _L, _P, _P, _P, _S, _P]),
    "bulkx_stable_sort_by_key": (_I, [_I, _I, _I, _P, _P, _L, _P, _S, _P]),
    "bulkx_reduce_by_key": (_I, [_I, _I, _I, _I, _P, _P, _L, _P, _P, _P, _P, _S, _P]),
    "bulkx_reduce_intervals": (_I, [_I, _I, _P, _L, _L, _L, _P, _P]),
    "bulkx_tail_flags": (_I, [_I, _I, _P, _L, _P, _P]),
    "bulkx_head_flags": (_I, [_I, _I, _P, _L, _P, _P]),
    "bulkx_batch_create": (_I, [_I, _L, _P, ctypes.POINTER(_P)]),
    "bulkx_batch_destroy": (_I, [_P]),
    "bulkx_batch_set_jobs": (_I, [_P, ctypes.POINTER(Job), _I]),
    "bulkx_batch_launch": (_I, [_P, _I]),
    "bulkx_batch_launch_per_call": (_I, [_P]),
    "bulkx_batch_status": (_I, [_P]),
    "bulkx_batch_reserve": (_I, [_P, _L]),
    "bulkx_batch_scratch_bytes": (_I, [_P, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]),
    "bulkx_batch_last_timing": (_I, [_P, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
    "bulkx_batch_stage_timing": (_I, [_P, ctypes.POINTER(ctypes.c_float), _I]),
    "bulkx_batch_sort_geometry": (_I, [ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    "bulkx_batch_family_shapes": (_I, [ctypes.c_void_p, ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_int64)]),
    "bulkx_batch_prepare_graphs": (_I, [_P]),
    "bulkx_server_create": (_I, [_L, _I, _I, _L, _I, ctypes.POINTER(_P)]),
    "bulkx_server_destroy": (_I, [_P]),
    "bulkx_server_submit": (_I, [_P, ctypes.POINTER(Job), _L, ctypes.POINTER(_L)]),
    "bulkx_server_submit_at": (_I, [_P, ctypes.POINTER(Job), _L, ctypes.c_double, ctypes.POINTER(_L)]),
    "bulkx_server_submit_host": (_I, [_P, ctypes.POINTER(Job), ctypes.c_double, ctypes.POINTER(_L),
                                      ctypes.POINTER(_L)]),
    "bulkx_server_now_us": (ctypes.c_double, [_P]),
    "bulkx_server_drain": (_I, [_P]),
    "bulkx_server_request_times": (_I, [_P, _L, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_double)]),
    "bulkx_server_stats": (_I, [_P, ctypes.POINTER(_L), ctypes.POINTER(_L), ctypes.POINTER(_L)]),
}

_lib = None


def header_symbols() -> list[str]:
    """Every function name declared in include/bulkx.h."""
    text = open(HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bulkx_[a-z0-9_]+)\s*\(", text)))


def lib() -> ctypes.CDLL:
    """The loaded native library (loaded once, with every signature bound)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise BulkLibraryMissing(
                f"{LIB_PATH} is not built; run `python -m bulk_amd.build` (or "
                "__graft_entry__.build()) first. There is no CPU fallback.")
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(l, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def check(status: int, where: str) -> None:
    if status != 0:
        detail = ""
        try:
            detail = lib().bulkx_last_error().decode() if status == 6 else ""
        except Exception:  # pragma: no cover - best effort detail
            pass
        raise BulkError(status, where, detail)
