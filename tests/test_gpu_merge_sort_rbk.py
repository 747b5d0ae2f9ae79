"""GPU parity: merge, merge_by_key, stable_sort_by_key, reduce_by_key and the
flag kernels through the C-ABI against the CPU oracle.  Keys, payloads, run
counts and integer sums are bit-exact; float32 sums are within
4*ceil(log2 n)*2^-24*sum|x| of the long-double oracle."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

MINSTD_SIZES = [48271, 153570, 597830, 20861, 391409]


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _host(t):
    return t.cpu().numpy()


# ----------------------------------------------------------------- merge
@pytest.mark.parametrize("n1,n2", [(0, 1), (1, 0), (1, 1), (5, 3), (4095, 1), (4096, 4096),
                                   (10000, 3), (3, 10000), (123457, 65537)])
def test_merge_int32_bit_exact(bx, orc, n1, n2):
    print(f"merge n1={n1} n2={n2}", flush=True)
    a = np.sort(orc.gen_uniform_i32(1, -50, 50, n1))
    b = np.sort(orc.gen_uniform_i32(2, -50, 50, n2))
    out = _host(bx.my_merge(_dev(a), _dev(b)))
    assert np.array_equal(out, orc.merge(a, b))


def test_merge_reference_validate_shape(bx, orc):
    # merge.cu:258-278: hash(i) keys, sorted halves (n scaled down from 123456789)
    n = 1 << 22
    a = np.sort(orc.ref_hash_i32(n // 2))
    b = np.sort(orc.ref_hash_i32(n)[n // 2:])
    out = _host(bx.my_merge(_dev(a), _dev(b)))
    assert np.array_equal(out, orc.merge(a, b))


@pytest.mark.parametrize("dt", [np.uint32, np.int64, np.float64, np.float32])
def test_merge_other_key_types(bx, orc, dt):
    if dt == np.uint32 and not hasattr(torch, "uint32"):
        pytest.skip("torch lacks uint32")
    a = np.sort((orc.gen_normal(3, 70001) * 100).astype(np.int64).astype(dt))
    b = np.sort((orc.gen_normal(4, 50003) * 100).astype(np.int64).astype(dt))
    out = _host(bx.my_merge(_dev(a), _dev(b)))
    assert np.array_equal(out, orc.merge(a, b))


def test_merge_greater(bx, orc):
    a = np.sort(orc.gen_uniform_i32(5, 0, 99, 30000))[::-1].copy()
    b = np.sort(orc.gen_uniform_i32(6, 0, 99, 20000))[::-1].copy()
    out = _host(bx.my_merge(_dev(a), _dev(b), comp="greater"))
    assert np.array_equal(out, orc.merge(a, b, comp="greater"))


@pytest.mark.parametrize("n", [1 << 12, (1 << 20) + 7])
def test_merge_by_key_ties_first_range_wins(bx, orc, n):
    # float keys with 10% drawn from a shared 1024-value pool; payload marks the source
    pool = orc.gen_uniform_f32(9, 1024)
    a = orc.gen_merge_keys(0, 0, n, pool, 0.1)
    b = orc.gen_merge_keys(0, 1, n + 3, pool, 0.1)
    va = np.arange(a.size, dtype=np.int32)
    vb = (np.int32(1 << 30) + np.arange(b.size)).astype(np.int32)
    ko, vo = bx.merge_by_key(_dev(a), _dev(va), _dev(b), _dev(vb))
    rk, rv = orc.merge_by_key(a, va, b, vb)
    assert np.array_equal(_host(ko), rk) and np.array_equal(_host(vo), rv)


def test_merge_by_key_8byte_payload(bx, orc):
    a = np.sort(orc.gen_uniform_i32(1, 0, 1000, 77777))
    b = np.sort(orc.gen_uniform_i32(2, 0, 1000, 55555))
    va = np.arange(a.size, dtype=np.int64) * 3
    vb = -np.arange(b.size, dtype=np.int64)
    ko, vo = bx.merge_by_key(_dev(a), _dev(va), _dev(b), _dev(vb))
    cat = np.concatenate([a, b])
    order = np.argsort(cat, kind="stable")
    assert np.array_equal(_host(ko), cat[order])
    assert np.array_equal(_host(vo), np.concatenate([va, vb])[order])


# ------------------------------------------------------------------ sort
@pytest.mark.parametrize("n", [1, 2, 15, 16, 17, 4095, 4096, 4097, 8193, 20000, 65537]
                         + MINSTD_SIZES)
def test_stable_sort_by_key_reference_inputs(bx, orc, n):
    # merge_sort_by_key.cu:315-362: keys hash(i), payload = index (stability)
    print(f"sort n={n}", flush=True)
    keys = orc.ref_hash_i32(n)
    vals = np.arange(n, dtype=np.int32)
    dk, dv = _dev(keys), _dev(vals)
    bx.stable_merge_sort_by_key(dk, dv)
    rk, rv = orc.stable_sort_by_key(keys, vals)
    assert np.array_equal(_host(dk), rk) and np.array_equal(_host(dv), rv)


@pytest.mark.parametrize("comp", ["less", "greater"])
def test_stable_sort_many_ties_is_stable(bx, orc, comp):
    n = 300001
    keys = (orc.gen_uniform_i32(7, 0, 9, n)).astype(np.int32)
    vals = np.arange(n, dtype=np.int32)
    dk, dv = _dev(keys), _dev(vals)
    bx.stable_merge_sort_by_key(dk, dv, comp=comp)
    rk, rv = orc.stable_sort_by_key(keys, vals, comp=comp)
    assert np.array_equal(_host(dk), rk) and np.array_equal(_host(dv), rv)


@pytest.mark.parametrize("dt", [np.float32, np.int64, np.float64])
def test_stable_sort_other_key_types(bx, orc, dt):
    n = 100003
    keys = (orc.gen_normal(8, n) * 50).astype(dt)
    vals = np.arange(n, dtype=np.int64)
    dk, dv = _dev(keys), _dev(vals)
    bx.stable_merge_sort_by_key(dk, dv)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(_host(dk), keys[order]) and np.array_equal(_host(dv), vals[order])


@pytest.mark.parametrize("dist", ["f32 [0,1)", "i32 all equal", "i32 87% one value",
                                  "i32 presorted", "f32 [0,1) greater"])
def test_stable_sort_skewed_digits(bx, dist):
    """Inputs whose digits are mostly one value in some radix pass (the
    counting histograms see up to 64 lanes on one digit), over many tiles."""
    rng = np.random.default_rng(21)
    n = 1_000_003
    if dist.startswith("f32"):
        keys = rng.random(n, dtype=np.float32)
    elif dist == "i32 all equal":
        keys = np.full(n, -77, np.int32)
    elif dist == "i32 87% one value":
        keys = rng.integers(-(1 << 31), 1 << 31, size=n, dtype=np.int64).astype(np.int32)
        keys[rng.random(n) < 0.87] = 123456
    else:
        keys = np.sort(rng.integers(-(1 << 31), 1 << 31, size=n, dtype=np.int64).astype(np.int32))
    comp = "greater" if dist.endswith("greater") else "less"
    vals = np.arange(n, dtype=np.int32)
    dk, dv = _dev(keys), _dev(vals)
    bx.stable_merge_sort_by_key(dk, dv, comp=comp)
    order = np.argsort(-keys if comp == "greater" else keys, kind="stable")
    assert np.array_equal(_host(dk), keys[order]) and np.array_equal(_host(dv), vals[order])


def test_stable_sort_keys_only(bx, orc):
    keys = orc.ref_hash_i32(200000)
    dk = _dev(keys)
    bx.stable_merge_sort_by_key(dk, None)
    assert np.array_equal(_host(dk), np.sort(keys, kind="stable"))


def test_stable_sort_u32_full_range_index_payload(bx, orc):
    if not hasattr(torch, "uint32"):
        pytest.skip("torch lacks uint32")
    n = 1 << 22
    keys = orc.gen_uniform_u32(0, n)
    vals = np.arange(n, dtype=np.uint32)
    dk, dv = _dev(keys), _dev(vals)
    bx.stable_merge_sort_by_key(dk, dv)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(_host(dk), keys[order]) and np.array_equal(_host(dv), vals[order])


# --------------------------------------------------------- reduce_by_key
@pytest.mark.parametrize("n", [1, 2, 3, 4095, 4096, 4097, 19999, 20000, 20001, 1 << 20]
                         + MINSTD_SIZES)
def test_reduce_by_key_reference_inputs(bx, orc, n):
    # reduce_by_key.cu:364-402: keys and values hash(i) % 10 -> short random runs
    print(f"rbk n={n}", flush=True)
    keys = orc.ref_hash_i32(n, mod10=True)
    vals = orc.ref_hash_i32(n, mod10=True)
    kr, vr, m = bx.my_reduce_by_key(_dev(keys), _dev(vals))
    rk, rv = orc.reduce_by_key(keys, vals)
    assert m == rk.size
    assert np.array_equal(_host(kr)[:m], rk) and np.array_equal(_host(vr)[:m], rv)


def test_reduce_by_key_single_run_and_all_distinct(bx, orc):
    n = 1 << 18
    keys = np.full(n, 7, np.int32)
    vals = orc.gen_uniform_i64(1, -(1 << 20), 1 << 20, n)
    kr, vr, m = bx.my_reduce_by_key(_dev(keys), _dev(vals))
    assert m == 1 and _host(kr)[0] == 7 and _host(vr)[0] == vals.sum()
    keys = np.arange(n, dtype=np.int32)
    kr, vr, m = bx.my_reduce_by_key(_dev(keys), _dev(vals))
    assert m == n and np.array_equal(_host(vr), vals)


@pytest.mark.parametrize("dist", ["geometric", "zipf"])
def test_reduce_by_key_int64_values_exact(bx, orc, dist):
    n = 1 << 22
    keys, runs = (orc.gen_runs_geometric(3, n) if dist == "geometric"
                  else orc.gen_runs_zipf(3, n))
    vals = orc.gen_uniform_i64(3, -(1 << 20), 1 << 20, n)
    kr, vr, m = bx.my_reduce_by_key(_dev(keys), _dev(vals))
    assert m == runs
    rk, rv = orc.reduce_by_key(keys, vals, textbook=True)
    assert np.array_equal(_host(kr)[:m], rk) and np.array_equal(_host(vr)[:m], rv)


@pytest.mark.parametrize("dist", ["geometric", "zipf"])
def test_reduce_by_key_float32_values_within_bound(bx, orc, dist):
    n = 1 << 22
    keys, runs = (orc.gen_runs_geometric(5, n) if dist == "geometric"
                  else orc.gen_runs_zipf(5, n))
    vals = orc.gen_uniform_f32(5, n, -1.0, 1.0)
    kr, vr, m = bx.my_reduce_by_key(_dev(keys), _dev(vals))
    assert m == runs
    assert np.array_equal(_host(kr)[:m], np.arange(runs, dtype=np.int32))
    starts = np.concatenate([[0], np.flatnonzero(np.diff(keys)) + 1])
    exact = np.add.reduceat(vals.astype(np.float64), starts)  # float64 run sums
    absum = np.add.reduceat(np.abs(vals.astype(np.float64)), starts)
    lens = np.diff(np.concatenate([starts, [n]]))
    got = _host(vr)[:m].astype(np.float64)
    k = np.maximum(1, np.ceil(np.log2(np.maximum(lens, 2))))
    # float64 run sums are exact to ~1e-16 relative for these lengths
    assert np.all(np.abs(got - exact) <= 4 * k * 2.0 ** -24 * absum + 1e-12 * absum)


@pytest.mark.parametrize("op", ["min", "max"])
def test_reduce_by_key_min_max(bx, orc, op):
    n = 500000
    keys, _ = orc.gen_runs_geometric(2, n, 30.0)
    vals = orc.gen_uniform_i32(2, -1000, 1000, n)
    kr, vr, m = bx.my_reduce_by_key(_dev(keys), _dev(vals), op=op)
    rk, rv = orc.reduce_by_key(keys, vals, op=op, textbook=True)
    assert m == rk.size and np.array_equal(_host(vr)[:m], rv)


def test_flags(bx, orc):
    keys = orc.ref_hash_i32(100000, mod10=True)
    assert np.array_equal(_host(bx.tail_flags(_dev(keys))), orc.tail_flags(keys))
    assert np.array_equal(_host(bx.head_flags(_dev(keys))), orc.head_flags(keys))


# ------------------------------------------------------- signed zeros
def _zeros_keys(seed, n):
    """float32 keys from {-1, -0.0, +0.0, 1}: -0.0 and +0.0 compare equal
    under less and equal_to, so only the input order separates them."""
    pool = np.array([-1.0, -0.0, 0.0, 1.0], dtype=np.float32)
    return pool[np.random.default_rng(seed).integers(0, 4, n)]


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def test_signed_zero_keys_sort_merge_reduce_by_key(bx, orc):
    """-0.0 and +0.0 are equivalent keys: stable sort and merge keep their
    input order (bits compared), and reduce_by_key puts them in one run whose
    output key is the run's last key (reduce_by_key.hpp:87-92)."""
    n = 100003
    k = _zeros_keys(1, n)
    v = np.arange(n, dtype=np.int32)
    dk, dv = _dev(k), _dev(v)
    bx.stable_merge_sort_by_key(dk, dv)
    rk, rv = orc.stable_sort_by_key(k.copy(), v.copy())
    assert np.array_equal(_bits(_host(dk)), _bits(rk)) and np.array_equal(_host(dv), rv)

    a = np.sort(_zeros_keys(2, 70001), kind="stable")
    b = np.sort(_zeros_keys(3, 50003), kind="stable")
    va = np.arange(a.size, dtype=np.int32)
    vb = (1 << 28) + np.arange(b.size, dtype=np.int32)
    ko, vo = bx.merge_by_key(_dev(a), _dev(va), _dev(b), _dev(vb))
    mk, mv = orc.merge_by_key(a, va, b, vb)
    assert np.array_equal(_bits(_host(ko)), _bits(mk)) and np.array_equal(_host(vo), mv)

    # runs of equal keys built from +-0.0 and +-1 in random lengths
    runs = np.random.default_rng(4).integers(1, 9, 40000)
    rkeys = np.repeat(np.where(np.arange(runs.size) % 2 == 0, 0.0, 1.0).astype(np.float32), runs)
    signs = np.where(np.random.default_rng(5).integers(0, 2, rkeys.size) == 0, -1.0, 1.0)
    rkeys = (rkeys * signs.astype(np.float32)).astype(np.float32)
    # adjacent runs of +-1 with different signs are different keys; +-0.0 are one key
    vals = np.random.default_rng(6).integers(-100, 100, rkeys.size).astype(np.int32)
    ko, vo, m = bx.my_reduce_by_key(_dev(rkeys), _dev(vals))
    ek, ev = orc.reduce_by_key(rkeys, vals)
    assert m == ek.size
    assert np.array_equal(_bits(_host(ko)[:m]), _bits(ek)) and np.array_equal(_host(vo)[:m], ev)
