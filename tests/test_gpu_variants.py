"""Every compiled tile shape (tuning variant) of every algorithm gives the
oracle's result: integer work, merge, sort and reduce_by_key bit-exact,
float sums within 4*ceil(log2 n)*eps*sum|x|.  Each test forces one variant
for every size class with bulkx_set_tuning and restores the packaged table
afterwards."""
import ctypes

import numpy as np
import pytest
import torch

from bulk_amd import _capi as C

pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _variants(algo):
    c = ctypes.c_int()
    C.check(C.lib().bulkx_tuning_variants(algo, ctypes.byref(c)), "variants")
    return c.value


@pytest.fixture
def force(bx):
    def f(algo, dtype, v):
        for sc in range(5):
            C.check(C.lib().bulkx_set_tuning(algo, dtype, sc, v), "bulkx_set_tuning")
    yield f
    C.check(C.lib().bulkx_load_tuning(None, None), "bulkx_load_tuning")


SIZES = [1, 4095, 4096, 4097, 8193, 16385, 100003, 1 << 20]


@pytest.mark.parametrize("v", [0, 1, 2, 3, 4])
def test_scan_variants(force, orc, v):
    """Variant 4 (65536-element tiles for 4-byte integers; float64 runs the
    default shape under it)."""
    from bulk_amd import api
    assert _variants(C.ALGO_INCLUSIVE_SCAN) == 5
    force(C.ALGO_INCLUSIVE_SCAN, C.INT32, v)
    force(C.ALGO_EXCLUSIVE_SCAN, C.INT32, v)
    force(C.ALGO_INCLUSIVE_SCAN, C.FLOAT64, v)
    for n in SIZES + [65535, 65536, 65537, 3 * 65536 + 1, (1 << 20) + 12345]:
        x = orc.gen_uniform_i32(n, -1000, 1000, n)
        d = _dev(x)
        assert np.array_equal(api.inclusive_scan(d, init=3).cpu().numpy(), orc.inclusive_scan(x, init=3)), n
        assert np.array_equal(api.exclusive_scan(d, init=3).cpu().numpy(), orc.exclusive_scan(x, init=3)), n
        f = orc.gen_uniform_f64(n, n)
        out = api.inclusive_scan(_dev(f)).cpu().numpy()
        hi, lo, ab = orc.prefix_sum_exact(f)
        assert np.all(np.abs(out - hi - lo) <= orc.float_tolerance(n, 2.0 ** -53, ab)), n


@pytest.mark.parametrize("v", [0, 1, 2])
def test_reduce_variants(force, orc, v):
    from bulk_amd import api
    force(C.ALGO_REDUCE, C.INT32, v)
    force(C.ALGO_REDUCE, C.FLOAT32, v)
    for n in SIZES + [(1 << 22) + 3]:
        x = orc.gen_uniform_i32(n, -1000, 1000, n)
        assert api.my_reduce(_dev(x), 5) == int(orc.reduce(x, 5)), n
        f = orc.gen_uniform_f32(n, n)
        hi, lo, ab = orc.sum_exact(f)
        assert abs(api.my_reduce(_dev(f), 0.0) - (hi + lo)) <= orc.float_tolerance(n, 2.0 ** -24, ab), n


@pytest.mark.parametrize("v", [0, 1, 2, 3, 4])
def test_merge_and_sort_variants(force, orc, v):
    from bulk_amd import api
    assert _variants(C.ALGO_SORT_BY_KEY) == 5 and _variants(C.ALGO_MERGE) == 2
    for a in (C.ALGO_MERGE, C.ALGO_MERGE_BY_KEY):
        force(a, C.FLOAT32, min(v, 1))
        force(a, C.INT32, min(v, 1))
    force(C.ALGO_SORT_BY_KEY, C.INT32, v)
    force(C.ALGO_SORT_BY_KEY, C.INT64, v)
    pool = orc.gen_uniform_f32(7, 64)
    for n1, n2 in [(1, 0), (4095, 1), (4096, 4097), (70000, 30001), (1 << 19, (1 << 19) + 7)]:
        a = orc.gen_merge_keys(n1, 0, n1, pool, 0.1)
        b = orc.gen_merge_keys(n1, 1, n2, pool, 0.1)
        va = np.arange(n1, dtype=np.int32)
        vb = (1 << 28) + np.arange(n2, dtype=np.int32)
        ko, vo = api.merge_by_key(_dev(a), _dev(va), _dev(b), _dev(vb))
        rk, rv = orc.merge_by_key(a, va, b, vb)
        assert np.array_equal(ko.cpu().numpy(), rk) and np.array_equal(vo.cpu().numpy(), rv), (n1, n2)
        assert np.array_equal(api.my_merge(_dev(a), _dev(b)).cpu().numpy(), orc.merge(a, b)), (n1, n2)
    for n in SIZES + [16384 * 5 + 3, 3 * (1 << 18) + 1]:
        k = orc.gen_uniform_i32(n, 0, max(2, n // 8), n)
        vals = np.arange(n, dtype=np.int32)
        dk, dv = _dev(k), _dev(vals)
        api.stable_merge_sort_by_key(dk, dv)
        order = np.argsort(k, kind="stable")
        assert np.array_equal(dk.cpu().numpy(), k[order]) and np.array_equal(dv.cpu().numpy(), vals[order]), n
        k8 = orc.gen_uniform_i64(n, -50, 50, n)
        d8 = _dev(k8)
        api.stable_merge_sort_by_key(d8)
        assert np.array_equal(d8.cpu().numpy(), np.sort(k8, kind="stable")), n


@pytest.mark.parametrize("v", list(range(12)))
def test_reduce_by_key_variants(force, orc, v):
    from bulk_amd import api
    for code in (C.INT32, C.INT64, C.FLOAT32):
        force(C.ALGO_REDUCE_BY_KEY, code, v)
    for n in SIZES + [(1 << 21) + 5]:
        keys, _ = orc.gen_runs_geometric(n, n, 8.0)
        for vals in (orc.gen_uniform_i32(n + 1, -50, 50, n), orc.gen_uniform_i64(n + 2, -(1 << 20), 1 << 20, n)):
            ko, vo, m = api.my_reduce_by_key(_dev(keys), _dev(vals))
            rk, rv = orc.reduce_by_key(keys, vals, textbook=True)
            assert m == rk.size and np.array_equal(ko[:m].cpu().numpy(), rk) and \
                np.array_equal(vo[:m].cpu().numpy(), rv), (n, vals.dtype)
        f = orc.gen_uniform_f32(n + 3, n, -1.0, 1.0)
        ko, vo, m = api.my_reduce_by_key(_dev(keys), _dev(f))
        rk, _ = orc.reduce_by_key(keys, f.astype(np.float64), textbook=True)
        assert m == rk.size and np.array_equal(ko[:m].cpu().numpy(), rk), n
    # runs longer than a tile: tiles with no head carry the open run (and,
    # in the three-pass form, tiles where no run ends are skipped: whole
    # tiles with 8-byte values, the value loads with 4-byte ones)
    for n, run in ((300001, 50000), (1 << 20, 1 << 20)):
        keys = (np.arange(n) // run).astype(np.int32)
        for vals in (orc.gen_uniform_i64(n + 7, -(1 << 20), 1 << 20, n),
                     orc.gen_uniform_i32(n + 8, -(1 << 10), 1 << 10, n)):
            ko, vo, m = api.my_reduce_by_key(_dev(keys), _dev(vals))
            rk, rv = orc.reduce_by_key(keys, vals, textbook=True)
            assert m == rk.size and np.array_equal(ko[:m].cpu().numpy(), rk) and \
                np.array_equal(vo[:m].cpu().numpy(), rv), (n, vals.dtype)
        f = orc.gen_uniform_f32(n + 9, n, -1.0, 1.0).astype(np.float64)
        ko, vo, m = api.my_reduce_by_key(_dev(keys), _dev(f.astype(np.float32)))
        starts = np.flatnonzero(np.r_[True, keys[1:] != keys[:-1]])
        exact = np.add.reduceat(f, starts)
        absum = np.add.reduceat(np.abs(f), starts)
        lens = np.diff(np.r_[starts, n])
        # the float tolerance 4 * ceil(log2 len) * 2^-24 * sum|x| per run
        bound = 4 * np.maximum(1, np.ceil(np.log2(np.maximum(lens, 2)))) * 2.0 ** -24 * absum
        assert m == starts.size
        assert np.all(np.abs(vo[:m].cpu().numpy().astype(np.float64) - exact) <= bound), n


@pytest.mark.parametrize("v", list(range(12)))
def test_reduce_by_key_last_run_key_zero(force, orc, v):
    """A run ending the array whose key equals the loads' fill value (0),
    with the array's end on a lane boundary inside a partial chunk and
    elsewhere: the last run is still written and counted."""
    from bulk_amd import api
    for code in (C.INT32, C.INT64, C.FLOAT32):
        force(C.ALGO_REDUCE_BY_KEY, code, v)
    for n in (16, 100, 70000, 16 * 1000, 4096 * 3 + 16, 8192 + 32, 8192 * 5 + 16 * 23):
        for keys in (np.zeros(n, np.int32), np.r_[np.full(n // 2, 5), np.zeros(n - n // 2)].astype(np.int32)):
            vals = orc.gen_uniform_i64(n, -(1 << 20), 1 << 20, n)
            ko, vo, m = api.my_reduce_by_key(_dev(keys), _dev(vals))
            rk, rv = orc.reduce_by_key(keys, vals, textbook=True)
            assert m == rk.size and np.array_equal(ko[:m].cpu().numpy(), rk) and \
                np.array_equal(vo[:m].cpu().numpy(), rv), (n, v)


RADIX_SIZES = [2, 8191, 8192, 8193, 16383, 16384, 16385, 3 * 8192 + 17, 100003, (1 << 20) + 5,
               (1 << 22) + 3]


@pytest.mark.parametrize("kt", ["int32", "uint32", "float32"])
@pytest.mark.parametrize("payload", ["none", "u32", "u64"])
def test_radix_sort_variant(force, orc, kt, payload):
    """Sort variant 4 (the radix passes) equals a stable argsort bit for bit:
    keys of every 4-byte type, keys only or with 4- or 8-byte payloads,
    less and greater, many ties, and -0.0 / +0.0 as one key."""
    from bulk_amd import api
    code = {"int32": C.INT32, "uint32": C.UINT32, "float32": C.FLOAT32}[kt]
    force(C.ALGO_SORT_BY_KEY, code, 4)
    tdt = {"int32": torch.int32, "uint32": torch.uint32, "float32": torch.float32}[kt]
    for n in RADIX_SIZES:
        if kt == "uint32":
            k = orc.gen_uniform_u32(n, n)
        elif kt == "int32":
            k = orc.gen_uniform_i32(n, -max(2, n // 8), max(2, n // 8), n)  # ties
        else:
            pool = np.array([-1.5, -0.0, 0.0, 2.0, 1e-30, -np.inf, np.inf], np.float32)
            k = np.where(orc.gen_uniform_u32(n + 1, n) % 3 == 0,
                         pool[orc.gen_uniform_u32(n + 2, n) % pool.size],
                         orc.gen_normal(n, n, np.float32))
            k = k.astype(np.float32)
        for comp in ("less", "greater"):
            dk = _dev(k.view(np.int32)).view(tdt)
            idx = np.arange(n, dtype=np.uint64)
            dv = None
            if payload == "u32":
                dv = _dev(idx.astype(np.int32))
            elif payload == "u64":
                dv = _dev((idx * 3 + (1 << 40)).astype(np.int64))
            api.stable_merge_sort_by_key(dk, dv, comp=comp)
            if kt == "float32":
                key = k.astype(np.float64)  # -0.0 == 0.0 under both orders
            else:
                key = k.astype(np.int64)
            order = np.argsort(key if comp == "less" else -key, kind="stable")
            got = dk.view(torch.int32).cpu().numpy().view(k.dtype)
            assert np.array_equal(got.view(np.uint32), k[order].view(np.uint32)), (n, comp)
            if payload == "u32":
                assert np.array_equal(dv.cpu().numpy(), order.astype(np.int32)), (n, comp)
            elif payload == "u64":
                assert np.array_equal(dv.cpu().numpy(), (order.astype(np.int64) * 3 + (1 << 40))), (n, comp)
