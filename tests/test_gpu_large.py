"""Inputs past 2^31 elements, the limit of the reference's `int` size_type
(bulk/execution_policy.hpp:66,109): scan, reduce, reduce_by_key, merge and
stable_sort_by_key at n > 2^31 through the C-ABI.  Outputs are checked chunk
by chunk on the GPU against closed forms (the inputs are periodic or
arithmetic sequences), so nothing of this size is copied to the host."""
import pytest
import torch

from bulk_amd import api

pytestmark = pytest.mark.gpu

CHUNK = 1 << 27


def _need_gib(gib):
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    free, _ = torch.cuda.mem_get_info()
    if free < gib * (1 << 30):
        pytest.skip(f"needs {gib} GiB of free device memory")


def _chunks(n):
    for s in range(0, n, CHUNK):
        yield s, min(n, s + CHUNK)


def _idx(s, e):
    return torch.arange(s, e, dtype=torch.int64, device="cuda")


def _u32_bits(v):
    """int64 values in [0, 2^32) -> int32 tensor with the same bits."""
    return torch.where(v >= (1 << 31), v - (1 << 32), v).to(torch.int32)


def _u32_read(t, s, e):
    return t.view(torch.int32)[s:e].to(torch.int64) & 0xFFFFFFFF


@pytest.fixture(autouse=True)
def _free_after():
    yield
    if torch.cuda.is_available():
        torch.cuda.synchronize()
        torch.cuda.empty_cache()


def test_scan_and_reduce_past_2_31():
    """x[i] = (i mod 7) - 3: inclusive prefix r(r+1)/2 - 3(r+1) with r = i mod 7."""
    n = (1 << 31) + 4097
    _need_gib(20)
    x = torch.empty(n, dtype=torch.int32, device="cuda")
    for s, e in _chunks(n):
        x[s:e] = (_idx(s, e) % 7 - 3).to(torch.int32)
    y = api.inclusive_scan(x)
    bad = 0
    for s, e in _chunks(n):
        r = _idx(s, e) % 7
        bad += int((y[s:e].to(torch.int64) != r * (r + 1) // 2 - 3 * (r + 1)).sum())
    assert bad == 0
    r = (n - 1) % 7
    assert api.my_reduce(x, 13) == 13 + r * (r + 1) // 2 - 3 * (r + 1)


def test_reduce_by_key_past_2_31():
    """keys i // 5, values (i mod 5) + 1: every run sums to 15, the last
    (two elements) to 3; keys out are 0, 1, 2, ..."""
    n = (1 << 31) + 4099
    runs = (n + 4) // 5
    _need_gib(24)
    keys = torch.empty(n, dtype=torch.int32, device="cuda")
    vals = torch.empty(n, dtype=torch.int32, device="cuda")
    for s, e in _chunks(n):
        i = _idx(s, e)
        keys[s:e] = (i // 5).to(torch.int32)
        vals[s:e] = (i % 5 + 1).to(torch.int32)
    ko = torch.empty(runs, dtype=torch.int32, device="cuda")
    vo = torch.empty(runs, dtype=torch.int32, device="cuda")
    ko, vo, m = api.my_reduce_by_key(keys, vals, ko, vo)
    assert m == runs
    bad = 0
    for s, e in _chunks(runs):
        bad += int((ko[s:e].to(torch.int64) != _idx(s, e)).sum())
        want = torch.full((e - s,), 15, dtype=torch.int32, device="cuda")
        if e == runs:
            want[-1] = 3
        bad += int((vo[s:e] != want).sum())
    assert bad == 0


def test_merge_past_2_31():
    """a = 0, 2, 4, ... (n1), b = 1, 3, 5, ... (n2 < n1), uint32 keys:
    output p is p below 2*n2 and 2*(p - n2) after."""
    if not hasattr(torch, "uint32"):
        pytest.skip("torch has no uint32")
    n1, n2 = (1 << 30) + 1000, (1 << 30) + 24
    n = n1 + n2
    _need_gib(20)
    a = torch.empty(n1, dtype=torch.int32, device="cuda")
    b = torch.empty(n2, dtype=torch.int32, device="cuda")
    for s, e in _chunks(n1):
        a[s:e] = _u32_bits(2 * _idx(s, e))
    for s, e in _chunks(n2):
        b[s:e] = _u32_bits(2 * _idx(s, e) + 1)
    out = api.my_merge(a.view(torch.uint32), b.view(torch.uint32))
    bad = 0
    for s, e in _chunks(n):
        p = _idx(s, e)
        want = torch.where(p < 2 * n2, p, 2 * (p - n2))
        bad += int((_u32_read(out, s, e) != want).sum())
    assert bad == 0


def test_stable_sort_by_key_past_2_31():
    """keys (n-1-i) // 2 (pairs of equal keys in descending order), values i:
    position p holds key p // 2 and, stably, value n - 2 - 2*(p // 2) + p mod 2."""
    if not hasattr(torch, "uint32"):
        pytest.skip("torch has no uint32")
    n = (1 << 31) + 2
    _need_gib(64)
    keys = torch.empty(n, dtype=torch.int32, device="cuda")
    vals = torch.empty(n, dtype=torch.int32, device="cuda")
    for s, e in _chunks(n):
        i = _idx(s, e)
        keys[s:e] = _u32_bits((n - 1 - i) // 2)
        vals[s:e] = _u32_bits(i)
    api.stable_merge_sort_by_key(keys.view(torch.uint32), vals.view(torch.uint32))
    bad = 0
    for s, e in _chunks(n):
        p = _idx(s, e)
        bad += int((_u32_read(keys, s, e) != p // 2).sum())
        bad += int((_u32_read(vals, s, e) != n - 2 - 2 * (p // 2) + p % 2).sum())
    assert bad == 0
