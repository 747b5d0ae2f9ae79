"""Python host mirror of the reference's device-wide drivers.

Each function keeps the name, argument meaning and error behaviour of the
reference driver it stands for and calls the C-ABI of libbulkx.so on the
current torch stream.  Tensors must live on the GPU; torch only provides
device memory and the stream.  A missing library raises
``BulkLibraryMissing``; a failing call raises ``BulkError``.
"""
from __future__ import annotations

import ctypes

import torch

from . import _capi as C

_DTYPES = {
    torch.int32: C.INT32, torch.int64: C.INT64, torch.float32: C.FLOAT32,
    torch.float64: C.FLOAT64,
}
for _name, _code in (("uint32", C.UINT32), ("uint64", C.UINT64)):
    if hasattr(torch, _name):
        _DTYPES[getattr(torch, _name)] = _code

_OPS = {"plus": C.PLUS, "min": C.MIN, "max": C.MAX}
_CMPS = {"less": C.LESS, "greater": C.GREATER}
_PREDS = {"equal_to": C.EQUAL_TO}

_CTYPE = {C.INT32: ctypes.c_int32, C.UINT32: ctypes.c_uint32, C.INT64: ctypes.c_int64,
          C.UINT64: ctypes.c_uint64, C.FLOAT32: ctypes.c_float, C.FLOAT64: ctypes.c_double}


def dtype_code(t: torch.Tensor) -> int:
    try:
        return _DTYPES[t.dtype]
    except KeyError:
        raise C.BulkError(2, f"dtype {t.dtype}") from None


def _stream(stream=None) -> ctypes.c_void_p:
    s = torch.cuda.current_stream() if stream is None else stream
    return ctypes.c_void_p(s.cuda_stream)


def _ptr(t: torch.Tensor | None) -> ctypes.c_void_p:
    if t is None:
        return ctypes.c_void_p(0)
    if not t.is_cuda:
        raise C.BulkError(1, "tensor is not on the GPU")
    if not t.is_contiguous():
        raise C.BulkError(1, "tensor is not contiguous")
    return ctypes.c_void_p(t.data_ptr())


def _scalar(code: int, value):
    return _CTYPE[code](value)


class Workspace:
    """Caller-owned scratch so calls do not allocate (see bulkx_workspace_size)."""

    def __init__(self, nbytes: int, device=None):
        self.nbytes = int(nbytes)
        self.buf = torch.empty(max(self.nbytes, 1), dtype=torch.uint8,
                               device=device or torch.device("cuda"))

    @staticmethod
    def for_algo(algo: int, dtype: int, value_dtype: int, n1: int, n2: int = 0):
        nb = ctypes.c_size_t(0)
        C.check(C.lib().bulkx_workspace_size(algo, dtype, value_dtype, n1, n2, ctypes.byref(nb)),
                "bulkx_workspace_size")
        return Workspace(nb.value)

    def args(self):
        return ctypes.c_void_p(self.buf.data_ptr()), ctypes.c_size_t(self.nbytes)


def _ws(ws: Workspace | None):
    if ws is None:
        return ctypes.c_void_p(0), ctypes.c_size_t(0)
    return ws.args()


def inclusive_scan(first: torch.Tensor, result: torch.Tensor | None = None, init=None,
                   op: str = "plus", workspace: Workspace | None = None, stream=None):
    """::inclusive_scan(first, last, result, init, binary_op), scan.cu:87-148.
    Returns `result` (allocated when None; may be `first` for in-place)."""
    code = dtype_code(first)
    if result is None:
        result = torch.empty_like(first)
    if result.dtype != first.dtype or result.numel() < first.numel():
        raise C.BulkError(1, "inclusive_scan: result must match first")
    init_c = None if init is None else _scalar(code, init)
    w, wb = _ws(workspace)
    C.check(C.lib().bulkx_inclusive_scan(code, _OPS[op], _ptr(first), _ptr(result), first.numel(),
                                         None if init_c is None else ctypes.byref(init_c), w, wb,
                                         _stream(stream)), "bulkx_inclusive_scan")
    return result


def exclusive_scan(first: torch.Tensor, result: torch.Tensor | None = None, init=0,
                   op: str = "plus", workspace: Workspace | None = None, stream=None):
    """Device-wide bulk::exclusive_scan (scan.hpp:553-591)."""
    code = dtype_code(first)
    if result is None:
        result = torch.empty_like(first)
    init_c = _scalar(code, init)
    w, wb = _ws(workspace)
    C.check(C.lib().bulkx_exclusive_scan(code, _OPS[op], _ptr(first), _ptr(result), first.numel(),
                                         ctypes.byref(init_c), w, wb, _stream(stream)),
            "bulkx_exclusive_scan")
    return result


def reduce_into(first: torch.Tensor, out: torch.Tensor, init=0, op: str = "plus",
                workspace: Workspace | None = None, stream=None) -> torch.Tensor:
    """Asynchronous reduce: writes the result into out[0] on the device."""
    code = dtype_code(first)
    if out.dtype != first.dtype or out.numel() < 1:
        raise C.BulkError(1, "reduce: out must hold one element of the input dtype")
    init_c = _scalar(code, init)
    w, wb = _ws(workspace)
    C.check(C.lib().bulkx_reduce(code, _OPS[op], _ptr(first), first.numel(), ctypes.byref(init_c),
                                 _ptr(out), w, wb, _stream(stream)), "bulkx_reduce")
    return out


def my_reduce(first: torch.Tensor, init=0, op: str = "plus", workspace: Workspace | None = None):
    """T my_reduce(first, last, init, binary_op), reduce.cu:58-97: returns the
    value to the host (a blocking read, as reduce.cu:96 does)."""
    out = torch.empty(1, dtype=first.dtype, device=first.device)
    reduce_into(first, out, init, op, workspace)
    return out.item()


def my_merge(first1: torch.Tensor, first2: torch.Tensor, result: torch.Tensor | None = None,
             comp: str = "less", workspace: Workspace | None = None, stream=None):
    """my_merge(first1, last1, first2, last2, result, comp), merge.cu:127-162."""
    code = dtype_code(first1)
    if first2.dtype != first1.dtype:
        raise C.BulkError(1, "merge: key dtypes differ")
    if result is None:
        result = torch.empty(first1.numel() + first2.numel(), dtype=first1.dtype,
                             device=first1.device)
    w, wb = _ws(workspace)
    C.check(C.lib().bulkx_merge(code, _CMPS[comp], _ptr(first1), first1.numel(), _ptr(first2),
                                first2.numel(), _ptr(result), w, wb, _stream(stream)),
            "bulkx_merge")
    return result


def merge_by_key(keys1, values1, keys2, values2, keys_result=None, values_result=None,
                 comp: str = "less", workspace: Workspace | None = None, stream=None):
    """Device-wide bulk::merge_by_key (merge.hpp:507-607)."""
    kc, vc = dtype_code(keys1), dtype_code(values1)
    n = keys1.numel() + keys2.numel()
    if keys_result is None:
        keys_result = torch.empty(n, dtype=keys1.dtype, device=keys1.device)
    if values_result is None:
        values_result = torch.empty(n, dtype=values1.dtype, device=values1.device)
    w, wb = _ws(workspace)
    C.check(C.lib().bulkx_merge_by_key(kc, vc, _CMPS[comp], _ptr(keys1), _ptr(values1),
                                       keys1.numel(), _ptr(keys2), _ptr(values2), keys2.numel(),
                                       _ptr(keys_result), _ptr(values_result), w, wb,
                                       _stream(stream)), "bulkx_merge_by_key")
    return keys_result, values_result


def stable_merge_sort_by_key(keys: torch.Tensor, values: torch.Tensor | None = None,
                             comp: str = "less", workspace: Workspace | None = None,
                             stream=None):
    """stable_merge_sort_by_key(keys_first, keys_last, values_first, comp),
    merge_sort_by_key.cu:150-209 (in place)."""
    kc = dtype_code(keys)
    vc = -1 if values is None else dtype_code(values)
    w, wb = _ws(workspace)
    C.check(C.lib().bulkx_stable_sort_by_key(kc, vc, _CMPS[comp], _ptr(keys), _ptr(values),
                                             keys.numel(), w, wb, _stream(stream)),
            "bulkx_stable_sort_by_key")
    return keys, values


def reduce_by_key_async(keys, values, keys_result, values_result, num_runs, pred="equal_to",
                        op="plus", workspace: Workspace | None = None, stream=None):
    """Asynchronous reduce_by_key: the run count lands in num_runs[0] (int64, device)."""
    kc, vc = dtype_code(keys), dtype_code(values)
    w, wb = _ws(workspace)
    C.check(C.lib().bulkx_reduce_by_key(kc, vc, _PREDS[pred], _OPS[op], _ptr(keys), _ptr(values),
                                        keys.numel(), _ptr(keys_result), _ptr(values_result),
                                        _ptr(num_runs), w, wb, _stream(stream)),
            "bulkx_reduce_by_key")


def my_reduce_by_key(keys, values, keys_result=None, values_result=None, pred="equal_to",
                     op="plus", workspace: Workspace | None = None):
    """my_reduce_by_key(keys_first, keys_last, values_first, keys_result,
    values_result, pred, binary_op), reduce_by_key.cu:176-272.  Returns
    (keys_result, values_result, count) with the count read back like
    reduce_by_key.cu:269."""
    n = keys.numel()
    if keys_result is None:
        keys_result = torch.empty(max(n, 1), dtype=keys.dtype, device=keys.device)
    if values_result is None:
        values_result = torch.empty(max(n, 1), dtype=values.dtype, device=values.device)
    cnt = torch.zeros(1, dtype=torch.int64, device=keys.device)
    reduce_by_key_async(keys, values, keys_result, values_result, cnt, pred, op, workspace)
    m = int(cnt.item())
    check_stream()  # the call blocks anyway: a look-back give-up raises here
    return keys_result, values_result, m


def reduce_intervals(first: torch.Tensor, num_partitions: int, tile_size: int,
                     result: torch.Tensor | None = None, op: str = "plus", stream=None):
    """reduce_intervals(first, aligned_decomposition(n, P, tile), result, op),
    reduce_intervals.hpp:31-40."""
    code = dtype_code(first)
    if result is None:
        result = torch.empty(max(num_partitions, 1), dtype=first.dtype, device=first.device)
    C.check(C.lib().bulkx_reduce_intervals(code, _OPS[op], _ptr(first), first.numel(),
                                           num_partitions, tile_size, _ptr(result),
                                           _stream(stream)), "bulkx_reduce_intervals")
    return result


def tail_flags(keys: torch.Tensor, pred: str = "equal_to", stream=None) -> torch.Tensor:
    """Materialised tail_flags (tail_flags.hpp:26-109) as uint8."""
    out = torch.empty(keys.numel(), dtype=torch.uint8, device=keys.device)
    C.check(C.lib().bulkx_tail_flags(dtype_code(keys), _PREDS[pred], _ptr(keys), keys.numel(),
                                     _ptr(out), _stream(stream)), "bulkx_tail_flags")
    return out


def head_flags(keys: torch.Tensor, pred: str = "equal_to", stream=None) -> torch.Tensor:
    """Materialised head_flags (head_flags.hpp:26-217) as uint8."""
    out = torch.empty(keys.numel(), dtype=torch.uint8, device=keys.device)
    C.check(C.lib().bulkx_head_flags(dtype_code(keys), _PREDS[pred], _ptr(keys), keys.numel(),
                                     _ptr(out), _stream(stream)), "bulkx_head_flags")
    return out


def stream_status(stream=None) -> int:
    """Waits for `stream` and returns BULKX_ERR_LOOKBACK (11) if a look-back
    wait on this device gave up since the last query (its launch's results
    are wrong), else 0 (bulkx_stream_status)."""
    return int(C.lib().bulkx_stream_status(_stream(stream)))


def check_stream(stream=None) -> None:
    """stream_status, raised as BulkError when non-zero."""
    C.check(stream_status(stream), "bulkx_stream_status")


def set_lookback_spin_limit(polls: int) -> None:
    """Polls a look-back wait makes before giving up (0: give up at once;
    a knob for tests of the error path)."""
    C.check(C.lib().bulkx_set_lookback_spin_limit(int(polls)), "bulkx_set_lookback_spin_limit")


def device_info() -> dict:
    cus, lds, wave, mt = (ctypes.c_int() for _ in range(4))
    hbm = ctypes.c_longlong()
    C.check(C.lib().bulkx_device_info(ctypes.byref(cus), ctypes.byref(lds), ctypes.byref(wave),
                                      ctypes.byref(mt), ctypes.byref(hbm)), "bulkx_device_info")
    return {"num_cus": cus.value, "lds_bytes_per_cu": lds.value, "wavefront_size": wave.value,
            "max_threads_per_group": mt.value, "hbm_bytes": hbm.value}


def tuned_variant(algo: int, dtype: int, n: int) -> int:
    """The tile-shape variant the tuning table selects for (algo, dtype, n)."""
    v = ctypes.c_int()
    C.check(C.lib().bulkx_tuned_variant(algo, dtype, n, ctypes.byref(v)), "bulkx_tuned_variant")
    return v.value


def choose_sizes(algo: int, dtype: int, n: int) -> tuple[int, int, int]:
    """bulk::choose_sizes counterpart: (num_groups, group_size, grain_size)."""
    g, s, r = ctypes.c_int64(), ctypes.c_int(), ctypes.c_int()
    C.check(C.lib().bulkx_choose_sizes(algo, dtype, n, ctypes.byref(g), ctypes.byref(s),
                                       ctypes.byref(r)), "bulkx_choose_sizes")
    return g.value, s.value, r.value
