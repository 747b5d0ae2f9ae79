"""GPU test of bulkx_inclusive_scan_noalloc, the allocation-free scan: its
output equals bulkx_inclusive_scan's on the same input (int32 bit for bit;
float32 within 4*ceil(log2 n)*2^-24*sum|x| of an FP64 prefix, like the
plain scan), on two calls in a row, the second smaller, so a stale
look-back record of the first call would show; sizes beyond the static
scratch are refused before anything is enqueued."""
import ctypes

import numpy as np
import pytest
import torch

from bulk_amd import _capi as C

pytestmark = pytest.mark.gpu


def _noalloc(dtype, x, y, init=None):
    lib = C.lib()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    ip = ctypes.byref(init) if init is not None else None
    return lib.bulkx_inclusive_scan_noalloc(dtype, C.PLUS, ctypes.c_void_p(x.data_ptr()),
                                            ctypes.c_void_p(y.data_ptr()), x.numel(), ip, s)


def _plain(dtype, x, y, init=None):
    lib = C.lib()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    ip = ctypes.byref(init) if init is not None else None
    return lib.bulkx_inclusive_scan(dtype, C.PLUS, ctypes.c_void_p(x.data_ptr()),
                                    ctypes.c_void_p(y.data_ptr()), x.numel(), ip, None, 0, s)


def _ints(n, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    return torch.randint(-1000, 1001, (n,), generator=g, dtype=torch.int32).cuda()


@pytest.mark.parametrize("n1,n2", [(1 << 24, (1 << 20) + 12345), (3_000_017, 8191)])
def test_noalloc_int32_equals_plain_twice(bx, n1, n2):
    for n, seed in ((n1, 1), (n2, 2)):  # second call smaller than the first
        x = _ints(n, seed)
        y0 = torch.empty_like(x)
        y1 = torch.full_like(x, 7)
        init = ctypes.c_int32(13)
        assert _plain(C.INT32, x, y0, init) == 0
        assert _noalloc(C.INT32, x, y1, init) == 0
        torch.cuda.synchronize()
        assert torch.equal(y0, y1), f"n={n}"
        ref = (torch.cumsum(x.cpu().long(), 0) + 13).to(torch.int32)
        assert torch.equal(y1.cpu(), ref)
    assert bx.stream_status() == 0


def test_noalloc_float32_within_bound_twice(bx):
    for n, seed in ((1 << 23, 3), (500_001, 4)):
        g = torch.Generator(device="cpu").manual_seed(seed)
        x = torch.rand(n, generator=g, dtype=torch.float32).cuda()
        y0, y1 = torch.empty_like(x), torch.empty_like(x)
        assert _plain(C.FLOAT32, x, y0) == 0
        assert _noalloc(C.FLOAT32, x, y1) == 0
        torch.cuda.synchronize()
        ref = torch.cumsum(x.double(), 0)
        bound = 4 * int(np.ceil(np.log2(n))) * 2.0 ** -24 * torch.cumsum(x.double().abs(), 0)
        assert bool(((y1.double() - ref).abs() <= bound).all())
        assert bool(((y0.double() - ref).abs() <= bound).all())
    assert bx.stream_status() == 0


def test_noalloc_rejects_sizes_beyond_scratch(bx):
    x = torch.zeros(16, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    lib = C.lib()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    # the check precedes any launch: the small buffers are never read
    for dt, n_max in ((C.FLOAT64, 1 << 29), (C.INT64, 1 << 30), (C.FLOAT32, 1 << 31)):
        st = lib.bulkx_inclusive_scan_noalloc(dt, C.PLUS, ctypes.c_void_p(x.data_ptr()),
                                              ctypes.c_void_p(y.data_ptr()), n_max + 1, None, s)
        assert st == 4  # BULKX_ERR_WORKSPACE_TOO_SMALL
    assert lib.bulkx_inclusive_scan_noalloc(C.INT32, C.PLUS, None, None, 0, None, s) == 0
