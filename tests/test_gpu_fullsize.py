"""Every benchmark configuration at its stated size, on the GPU.

C1  inclusive_scan of 2^24 int32 in [-1000, 1000], seed 0: bit-exact
    against the oracle's sequential body (scan.hpp:35-78).
C2  reduce and inclusive_scan of 2^28 float32 / float64, uniform [0,1)
    (seed 0) and normal(0,1) (seed 1): every element within
    4*ceil(log2 n)*eps*sum|x| of the oracle's compensated long-double prefix
    (eps = 2^-24 / 2^-53); the measured worst error is printed.
C3  merge_by_key of two sorted 2^27 float32 arrays with 10% of the keys from
    a shared 1024-value pool and int32 payloads; stable_sort_by_key of 2^26
    full-range uint32 keys with payload = index.  Checked on the GPU by the
    properties that fix a stable merge / sort exactly: the output keys are
    nondecreasing, (key, payload) pairs are a permutation of the input pairs
    (payload is a permutation and key == input key at the payload), and equal
    keys keep input order (payloads increase; merge payloads number the
    first range before the second, merge.hpp:111).
C4  reduce_by_key of 2^28 int32 keys in runs (geometric mean 8; Zipf 1.2
    capped at 2^16; key = run index), float32 values in [-1, 1] and int64
    values in [-2^20, 2^20]: run count and keys exact, int64 sums exact
    against index_add, float32 sums within the bound of a float64 index_add
    (whose own error is below 2^-37 sum|x| per run).
"""
import math

import numpy as np
import pytest
import torch

from bulk_amd import _capi as C

pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _bound_factor(n, eps):
    return 4 * max(1, math.ceil(math.log2(max(n, 2)))) * eps


# ------------------------------------------------------------------ C1
def test_c1_int32_scan_2_24_bit_exact(bx, orc):
    x = orc.gen_uniform_i32(0, -1000, 1000, 1 << 24)
    y = bx.inclusive_scan(_dev(x)).cpu().numpy()
    assert np.array_equal(y, orc.inclusive_scan(x))
    assert bx.stream_status() == 0


# ------------------------------------------------------------------ C2
def _c2_input(orc, dt, dist, n):
    if dist == "uniform":
        return orc.gen_uniform_f32(0, n) if dt == "f32" else orc.gen_uniform_f64(0, n)
    return orc.gen_normal(1, n, np.float32 if dt == "f32" else np.float64)


@pytest.mark.parametrize("dt", ["f32", "f64"])
@pytest.mark.parametrize("dist", ["uniform", "normal"])
def test_c2_reduce_2_28_within_bound(bx, orc, dt, dist):
    n = 1 << 28
    x = _c2_input(orc, dt, dist, n)
    eps = 2.0 ** -24 if dt == "f32" else 2.0 ** -53
    got = bx.my_reduce(_dev(x), 0.0)
    hi, lo, ab = orc.sum_exact(x)
    err = abs((got - hi) - lo)
    bound = _bound_factor(n, eps) * ab
    print(f"reduce {dt} {dist} 2^28: err {err:.3e} bound {bound:.3e} ratio {err / bound:.3e}")
    assert err <= bound


@pytest.mark.parametrize("dt", ["f32", "f64"])
@pytest.mark.parametrize("dist", ["uniform", "normal"])
def test_c2_inclusive_scan_2_28_within_bound(bx, orc, dt, dist):
    n = 1 << 28
    x = _c2_input(orc, dt, dist, n)
    eps = 2.0 ** -24 if dt == "f32" else 2.0 ** -53
    d = _dev(x)
    y = bx.inclusive_scan(d)
    del d
    y = y.cpu().numpy().astype(np.float64)
    assert bx.stream_status() == 0
    hi, lo, ab = orc.prefix_sum_exact(x)
    del x
    y -= hi
    y -= lo
    np.abs(y, out=y)
    ab *= _bound_factor(n, eps)
    ok = bool(np.all(y <= ab))
    ratio = float(np.max(y / np.maximum(ab, 1e-300)))
    print(f"inclusive_scan {dt} {dist} 2^28: worst err/bound {ratio:.3e}")
    assert ok


# ------------------------------------------------------------------ C3
def test_c3_merge_by_key_2x2_27(bx, orc):
    h = 1 << 27
    pool = orc.gen_uniform_f32(2, 1024)
    a = orc.gen_merge_keys(3, 0, h, pool, 0.1)
    b = orc.gen_merge_keys(3, 1, h, pool, 0.1)
    da, db = _dev(a), _dev(b)
    va = torch.arange(h, dtype=torch.int32, device="cuda")
    vb = va + h
    ko, vo = bx.merge_by_key(da, va, db, vb)
    assert bx.stream_status() == 0
    src = torch.cat([da, db])
    del da, db, va, vb
    n = 2 * h
    assert bool((ko[1:] >= ko[:-1]).all()), "keys nondecreasing"
    seen = torch.zeros(n, dtype=torch.uint8, device="cuda")
    seen[vo.long()] = 1
    assert bool(seen.all()), "payloads are a permutation"
    del seen
    assert torch.equal(ko, src[vo.long()]), "each key travels with its payload"
    ties = ko[1:] == ko[:-1]
    assert bool((vo[1:][ties] > vo[:-1][ties]).all()), "ties: first range first, input order"
    print(f"merge 2x2^27: {int(ties.sum())} adjacent equal keys, stable")


def test_c3_stable_sort_by_key_2_26(bx, orc):
    n = 1 << 26
    k0 = orc.gen_uniform_u32(4, n)
    dk0 = _dev(k0.view(np.int32)).view(torch.uint32)
    k = dk0.clone()
    v = torch.arange(n, dtype=torch.int32, device="cuda").view(torch.uint32)
    bx.stable_merge_sort_by_key(k, v)
    assert bx.stream_status() == 0
    ki = k.view(torch.int32).long() & 0xFFFFFFFF  # compare as unsigned
    vi = v.view(torch.int32).long()
    assert bool((ki[1:] >= ki[:-1]).all()), "keys nondecreasing"
    seen = torch.zeros(n, dtype=torch.uint8, device="cuda")
    seen[vi] = 1
    assert bool(seen.all()), "payload is a permutation"
    del seen
    assert torch.equal(k.view(torch.int32), dk0.view(torch.int32)[vi]), "key == input key at payload"
    ties = ki[1:] == ki[:-1]
    assert bool((vi[1:][ties] > vi[:-1][ties]).all()), "equal keys keep input order"
    print(f"sort 2^26: {int(ties.sum())} adjacent equal keys, stable")


# ------------------------------------------------------------------ C4
@pytest.mark.parametrize("runs", ["geometric", "zipf"])
@pytest.mark.parametrize("vt", ["f32", "i64"])
def test_c4_reduce_by_key_2_28(bx, orc, runs, vt):
    n = 1 << 28
    if runs == "geometric":
        keys, _ = orc.gen_runs_geometric(5, n, 8.0)
    else:
        keys, _ = orc.gen_runs_zipf(5, n, 1.2, 1 << 16)
    R = int(keys[-1]) + 1
    if vt == "f32":
        vals = orc.gen_uniform_f32(6, n, -1.0, 1.0)
    else:
        vals = orc.gen_uniform_i64(6, -(1 << 20), 1 << 20, n)
    dk, dv = _dev(keys), _dev(vals)
    del keys, vals
    ko = torch.empty_like(dk)
    vo = torch.empty_like(dv)
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    bx.reduce_by_key_async(dk, dv, ko, vo, cnt)
    m = int(cnt.item())
    assert bx.stream_status() == 0
    assert m == R, "run count"
    assert torch.equal(ko[:m], torch.arange(R, dtype=torch.int32, device="cuda")), "output keys"
    idx = dk.long()
    if vt == "i64":
        ref = torch.zeros(R, dtype=torch.int64, device="cuda").index_add_(0, idx, dv)
        assert torch.equal(vo[:m], ref), "int64 sums exact"
        print(f"reduce_by_key {runs} i64 2^28: {R} runs, sums exact")
    else:
        v64 = dv.double()
        ref = torch.zeros(R, dtype=torch.float64, device="cuda").index_add_(0, idx, v64)
        absum = torch.zeros(R, dtype=torch.float64, device="cuda").index_add_(0, idx, v64.abs())
        lens = torch.bincount(idx, minlength=R).double()
        bound = 4 * torch.clamp(torch.ceil(torch.log2(torch.clamp(lens, min=2))), min=1) * 2.0 ** -24 * absum
        err = (vo[:m].double() - ref).abs()
        ratio = float((err / bound.clamp_min(1e-300)).max())
        print(f"reduce_by_key {runs} f32 2^28: {R} runs, worst err/bound {ratio:.3e}")
        assert bool((err <= bound).all())
