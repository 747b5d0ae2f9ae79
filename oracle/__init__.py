"""CPU oracle (TEST INFRASTRUCTURE ONLY).

numpy front end of oracle/liboracle.so, the C++ restatement of the
reference's sequential agent bodies and device-wide decompositions.  Only
tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py may use
it, and only as the checker or the timed CPU restatement -- never as part of
the product path.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_TAG = {np.dtype(np.int32): "i32", np.dtype(np.uint32): "u32", np.dtype(np.int64): "i64",
        np.dtype(np.uint64): "u64", np.dtype(np.float32): "f32", np.dtype(np.float64): "f64"}
_CT = {"i32": ctypes.c_int32, "u32": ctypes.c_uint32, "i64": ctypes.c_int64,
       "u64": ctypes.c_uint64, "f32": ctypes.c_float, "f64": ctypes.c_double}
OPS = {"plus": 0, "min": 1, "max": 2}
CMPS = {"less": 0, "greater": 1}

_lib = None


def build() -> str:
    """Compile liboracle.so with the host compiler (make)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _fn(name, restype=None):
    f = getattr(lib(), name)
    f.restype = restype
    return f


def _tag(a) -> str:
    return _TAG[np.dtype(a.dtype)]


# ---------------------------------------------------------------- algorithms
def inclusive_scan(x: np.ndarray, init=None, op: str = "plus") -> np.ndarray:
    x = np.ascontiguousarray(x)
    t = _tag(x)
    out = np.empty_like(x)
    f = _fn(f"orc_inclusive_scan_{t}")
    f.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    ic = None if init is None else _CT[t](init)
    f(_p(x), x.size, None if ic is None else ctypes.byref(ic), OPS[op], _p(out))
    return out


def exclusive_scan(x: np.ndarray, init=0, op: str = "plus") -> np.ndarray:
    x = np.ascontiguousarray(x)
    t = _tag(x)
    out = np.empty_like(x)
    f = _fn(f"orc_exclusive_scan_{t}")
    f.argtypes = [ctypes.c_void_p, ctypes.c_int64, _CT[t], ctypes.c_int, ctypes.c_void_p]
    f(_p(x), x.size, init, OPS[op], _p(out))
    return out


def reduce(x: np.ndarray, init=0, op: str = "plus"):
    x = np.ascontiguousarray(x)
    t = _tag(x)
    f = _fn(f"orc_reduce_{t}", _CT[t])
    f.argtypes = [ctypes.c_void_p, ctypes.c_int64, _CT[t], ctypes.c_int]
    return x.dtype.type(f(_p(x), x.size, init, OPS[op]))


def reduce_intervals(x: np.ndarray, num_partitions: int, tile: int, op: str = "plus"):
    x = np.ascontiguousarray(x)
    t = _tag(x)
    out = np.empty(num_partitions, dtype=x.dtype)
    f = _fn(f"orc_reduce_intervals_{t}")
    f.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int,
                  ctypes.c_void_p]
    f(_p(x), x.size, num_partitions, tile, OPS[op], _p(out))
    return out


def aligned_ranges(n: int, num_partitions: int, tile: int):
    first = np.empty(num_partitions, np.int64)
    last = np.empty(num_partitions, np.int64)
    f = _fn("orc_aligned_ranges")
    f.argtypes = [ctypes.c_int64] * 3 + [ctypes.c_void_p] * 2
    f(n, num_partitions, tile, _p(first), _p(last))
    return first, last


def merge(a: np.ndarray, b: np.ndarray, comp: str = "less") -> np.ndarray:
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    t = _tag(a)
    out = np.empty(a.size + b.size, dtype=a.dtype)
    f = _fn(f"orc_merge_{t}")
    f.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                  ctypes.c_void_p]
    f(_p(a), a.size, _p(b), b.size, CMPS[comp], _p(out))
    return out


def merge_path(a: np.ndarray, b: np.ndarray, diag: int, comp: str = "less") -> int:
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    t = _tag(a)
    f = _fn(f"orc_merge_path_{t}", ctypes.c_int64)
    f.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                  ctypes.c_int]
    return int(f(_p(a), a.size, _p(b), b.size, diag, CMPS[comp]))


def merge_by_key(k1, v1, k2, v2, comp: str = "less"):
    k1, v1, k2, v2 = (np.ascontiguousarray(x) for x in (k1, v1, k2, v2))
    kt, vt = _tag(k1), _tag(v1)
    ko = np.empty(k1.size + k2.size, k1.dtype)
    vo = np.empty(k1.size + k2.size, v1.dtype)
    f = _fn(f"orc_merge_by_key_{kt}_{vt}")
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                  ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    f(_p(k1), _p(v1), k1.size, _p(k2), _p(v2), k2.size, CMPS[comp], _p(ko), _p(vo))
    return ko, vo


def stable_sort_by_key(keys, values=None, comp: str = "less"):
    keys = np.array(keys, copy=True)
    kt = _tag(keys)
    if values is None:
        f = _fn(f"orc_stable_sort_{kt}")
        f.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int]
        f(_p(keys), keys.size, CMPS[comp])
        return keys, None
    values = np.array(values, copy=True)
    f = _fn(f"orc_stable_sort_by_key_{kt}_{_tag(values)}")
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int]
    f(_p(keys), _p(values), keys.size, CMPS[comp])
    return keys, values


def reduce_by_key(keys, values, op: str = "plus", sms: int = 256, textbook: bool = False):
    """The reference's interval algorithm (default) or the textbook loop."""
    keys, values = np.ascontiguousarray(keys), np.ascontiguousarray(values)
    kt, vt = _tag(keys), _tag(values)
    ko = np.empty(max(keys.size, 1), keys.dtype)
    vo = np.empty(max(keys.size, 1), values.dtype)
    if textbook:
        f = _fn(f"orc_textbook_reduce_by_key_{kt}_{vt}", ctypes.c_int64)
        f.argtypes = [ctypes.c_void_p] * 2 + [ctypes.c_int64] + [ctypes.c_void_p] * 2 + [ctypes.c_int]
        m = f(_p(keys), _p(values), keys.size, _p(ko), _p(vo), OPS[op])
    else:
        f = _fn(f"orc_reduce_by_key_{kt}_{vt}", ctypes.c_int64)
        f.argtypes = ([ctypes.c_void_p] * 2 + [ctypes.c_int64] + [ctypes.c_void_p] * 2
                      + [ctypes.c_int, ctypes.c_int64])
        m = f(_p(keys), _p(values), keys.size, _p(ko), _p(vo), OPS[op], sms)
    return ko[:m], vo[:m]


def tail_flags(keys: np.ndarray) -> np.ndarray:
    keys = np.ascontiguousarray(keys, dtype=np.int32)
    out = np.empty(keys.size, np.uint8)
    f = _fn("orc_tail_flags_i32")
    f.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
    f(_p(keys), keys.size, _p(out))
    return out


def head_flags(keys: np.ndarray) -> np.ndarray:
    keys = np.ascontiguousarray(keys, dtype=np.int32)
    out = np.empty(keys.size, np.uint8)
    f = _fn("orc_head_flags_i32")
    f.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
    f(_p(keys), keys.size, _p(out))
    return out


# ------------------------------------------------- extended-precision checks
def prefix_sum_exact(x: np.ndarray, init: float = 0.0):
    """(hi, lo, abs_prefix): long-double prefix sums as hi+lo and prefix sum|x|."""
    x = np.ascontiguousarray(x)
    n = x.size
    hi, lo, ab = np.empty(n), np.empty(n), np.empty(n)
    f = _fn("orc_prefix_sum_ld_" + _tag(x))
    f.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_double] + [ctypes.c_void_p] * 3
    f(_p(x), n, float(init), _p(hi), _p(lo), _p(ab))
    return hi, lo, ab


def sum_exact(x: np.ndarray):
    """(hi, lo, abs_sum) of x in long double."""
    x = np.ascontiguousarray(x)
    hi, lo, ab = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    f = _fn("orc_sum_ld_" + _tag(x))
    f.argtypes = [ctypes.c_void_p, ctypes.c_int64] + [ctypes.c_void_p] * 3
    f(_p(x), x.size, ctypes.byref(hi), ctypes.byref(lo), ctypes.byref(ab))
    return hi.value, lo.value, ab.value


def float_tolerance(n: int, eps: float, abs_sum):
    """The stated bound 4 * ceil(log2 n) * eps * sum|x| (n >= 2; n = 1 -> one eps)."""
    import math
    k = max(1, math.ceil(math.log2(max(n, 2))))
    return 4.0 * k * eps * np.asarray(abs_sum, dtype=np.float64)


# ---------------------------------------------------------------- generators
def gen_uniform_i32(seed: int, lo: int, hi: int, n: int) -> np.ndarray:
    out = np.empty(n, np.int32)
    f = _fn("orc_gen_uniform_i32")
    f.argtypes = [ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p]
    f(seed, lo, hi, n, _p(out))
    return out


def gen_uniform_i64(seed: int, lo: int, hi: int, n: int) -> np.ndarray:
    out = np.empty(n, np.int64)
    f = _fn("orc_gen_uniform_i64")
    f.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p]
    f(seed, lo, hi, n, _p(out))
    return out


def gen_uniform_u32(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.uint32)
    f = _fn("orc_gen_uniform_u32")
    f.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_void_p]
    f(seed, n, _p(out))
    return out


def gen_uniform_f32(seed: int, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
    out = np.empty(n, np.float32)
    f = _fn("orc_gen_uniform_f32")
    f.argtypes = [ctypes.c_uint64, ctypes.c_float, ctypes.c_float, ctypes.c_int64, ctypes.c_void_p]
    f(seed, lo, hi, n, _p(out))
    return out


def gen_uniform_f64(seed: int, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
    out = np.empty(n, np.float64)
    f = _fn("orc_gen_uniform_f64")
    f.argtypes = [ctypes.c_uint64, ctypes.c_double, ctypes.c_double, ctypes.c_int64,
                  ctypes.c_void_p]
    f(seed, lo, hi, n, _p(out))
    return out


def gen_normal(seed: int, n: int, dtype=np.float64) -> np.ndarray:
    out = np.empty(n, dtype)
    f = _fn("orc_gen_normal_" + _TAG[np.dtype(dtype)])
    f.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_void_p]
    f(seed, n, _p(out))
    return out


def gen_runs_geometric(seed: int, n: int, mean: float = 8.0):
    keys = np.empty(n, np.int32)
    f = _fn("orc_gen_runs_geometric", ctypes.c_int64)
    f.argtypes = [ctypes.c_uint64, ctypes.c_double, ctypes.c_int64, ctypes.c_void_p]
    runs = f(seed, mean, n, _p(keys))
    return keys, int(runs)


def gen_runs_zipf(seed: int, n: int, s: float = 1.2, cap: int = 1 << 16):
    keys = np.empty(n, np.int32)
    f = _fn("orc_gen_runs_zipf", ctypes.c_int64)
    f.argtypes = [ctypes.c_uint64, ctypes.c_double, ctypes.c_int64, ctypes.c_int64,
                  ctypes.c_void_p]
    runs = f(seed, s, cap, n, _p(keys))
    return keys, int(runs)


def gen_merge_keys(seed: int, stream: int, n: int, pool: np.ndarray, frac: float = 0.1):
    pool = np.ascontiguousarray(pool, np.float32)
    out = np.empty(n, np.float32)
    f = _fn("orc_gen_merge_keys")
    f.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int64, ctypes.c_void_p,
                  ctypes.c_int64, ctypes.c_double, ctypes.c_void_p]
    f(seed, stream, n, _p(pool), pool.size, frac, _p(out))
    return out


def ref_hash_i32(n: int, mod10: bool = False) -> np.ndarray:
    out = np.empty(n, np.int32)
    f = _fn("orc_ref_hash_i32")
    f.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_void_p]
    f(n, 1 if mod10 else 0, _p(out))
    return out


def minstd(count: int) -> np.ndarray:
    out = np.empty(count, np.uint32)
    f = _fn("orc_minstd")
    f.argtypes = [ctypes.c_int64, ctypes.c_void_p]
    f(count, _p(out))
    return out


# ------------------------------------------------------------- CPU baseline
def mt_inclusive_scan_i32(x: np.ndarray, init: int, threads: int) -> np.ndarray:
    x = np.ascontiguousarray(x, np.int32)
    out = np.empty_like(x)
    f = _fn("orc_mt_inclusive_scan_i32")
    f.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int]
    f(_p(x), x.size, init, _p(out), threads)
    return out


def hardware_concurrency() -> int:
    f = _fn("orc_hardware_concurrency", ctypes.c_uint)
    f.argtypes = []
    return int(f())
