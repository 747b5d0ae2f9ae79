"""Pins the CPU oracle (oracle/) before any GPU result is compared with it.

Each restatement is checked against something it cannot share a mistake
with: the known answers of the reference's own checks, closed forms,
numpy/textbook routines and brute force on small inputs.
"""
import numpy as np
import pytest

# The 20 sizes the reference's validate loops draw with
# thrust::default_random_engine (minstd_rand, seed 1) % 2^20
# (scan.cu:222-229, merge_sort_by_key.cu:367-380, reduce_by_key.cu:413-420).
MINSTD_SIZES = [48271, 153570, 597830, 20861, 391409, 508195, 703057, 127065, 452803, 748159,
                412963, 711535, 426155, 1040717, 261801, 1035267, 363049, 978482, 295311, 650191]


def test_minstd_sizes_match_reference_engine(orc):
    assert [int(v) % (1 << 20) for v in orc.minstd(20)] == MINSTD_SIZES


# ------------------------------------------------------------------ scan
@pytest.mark.parametrize("n", [1, 2, 3, 1000, 1 << 16, 20861])
def test_scan_known_answer_ones_plus_13(orc, n):
    # scan.cu:158-184: all-ones input, init 13 -> result[i] = 14 + i
    out = orc.inclusive_scan(np.ones(n, np.int32), init=13)
    assert np.array_equal(out, 14 + np.arange(n, dtype=np.int32))


@pytest.mark.parametrize("dt", [np.int32, np.uint32, np.int64, np.uint64])
def test_scan_integer_matches_wrapping_cumsum(orc, dt):
    rng = np.random.default_rng(1)
    info = np.iinfo(dt)
    x = rng.integers(info.min, info.max, size=5000, dtype=dt, endpoint=True)
    ref = np.cumsum(x.astype(np.uint64) if dt in (np.int64, np.uint64) else x.astype(np.int64))
    ref = (ref.astype(np.uint64) + np.uint64(7)).astype(dt)  # wraps mod 2^w, then init 7
    assert np.array_equal(orc.inclusive_scan(x, init=7), ref)
    assert np.array_equal(orc.inclusive_scan(x), np.cumsum(x, dtype=dt))


def test_scan_min_max_match_numpy_accumulate(orc):
    x = orc.gen_uniform_i32(3, -1000, 1000, 4000)
    assert np.array_equal(orc.inclusive_scan(x, op="min"), np.minimum.accumulate(x))
    assert np.array_equal(orc.inclusive_scan(x, op="max"), np.maximum.accumulate(x))
    assert np.array_equal(orc.inclusive_scan(x, init=-5000, op="max"),
                          np.maximum.accumulate(np.maximum(x, -5000)))


def test_exclusive_scan_is_shifted_inclusive(orc):
    x = orc.gen_uniform_i32(4, -50, 50, 777)
    inc = orc.inclusive_scan(x, init=13)
    exc = orc.exclusive_scan(x, init=13)
    assert exc[0] == 13 and np.array_equal(exc[1:], inc[:-1])


def test_scan_float_against_long_double(orc):
    x = orc.gen_uniform_f32(0, 100000)
    hi, lo, ab = orc.prefix_sum_exact(x)
    # numpy float64 cumsum of float32 inputs is within 1e-12 relative of the long-double sum
    assert np.allclose(np.cumsum(x.astype(np.float64)), hi, rtol=1e-12, atol=0)
    # the sequential float32 body obeys the textbook recursive-summation bound
    # |err_i| <= i * u * sum_{j<=i} |x_j| (Higham, Accuracy and Stability, 4.2),
    # and it is NOT inside the 4*ceil(log2 n)*u bound for long inputs, which is
    # why the GPU path accumulates float32 in float64.
    seq = orc.inclusive_scan(x).astype(np.float64)
    i = np.arange(1, x.size + 1, dtype=np.float64)
    assert np.all(np.abs(seq - hi) <= i * 2.0 ** -24 * ab)
    assert np.any(np.abs(seq - hi) > orc.float_tolerance(x.size, 2.0 ** -24, ab))
    assert np.all(ab >= np.abs(hi))


def test_prefix_sum_exact_f64_brute_force(orc):
    from fractions import Fraction
    x = orc.gen_normal(5, 200)
    hi, lo, ab = orc.prefix_sum_exact(x)
    acc = Fraction(0)
    for i, v in enumerate(x):
        acc += Fraction(float(v))
        err = abs(Fraction(float(hi[i])) + Fraction(float(lo[i])) - acc)
        assert err <= Fraction(1, 2 ** 60) * (1 + abs(acc))


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_prefix_sum_exact_compensated(orc, dt):
    """The prefix oracle keeps the rounding error of every addition
    (Neumaier): [2^70, 1, -2^70] has prefixes 2^70, 2^70 + 1, 1, where plain
    long-double accumulation (64-bit mantissa) loses the 1 and ends at 0."""
    x = np.array([2.0 ** 70, 1.0, -(2.0 ** 70), 0.5, 2.0 ** -30], dtype=dt)
    hi, lo, ab = orc.prefix_sum_exact(x)
    assert hi[2] + lo[2] == 1.0
    assert hi[1] + lo[1] == 2.0 ** 70  # 2^70 + 1 rounded once to the long double grid
    assert hi[4] + lo[4] == 1.5 + 2.0 ** -30


@pytest.mark.parametrize("seed,n", [(21, 300000), (22, 1 << 20)])
def test_prefix_sum_exact_matches_fsum_at_samples(orc, seed, n):
    """At sampled positions the compensated prefix equals math.fsum of the
    slice (the exact sum rounded once) to within 2^-60 of sum|x|, on data
    with heavy cancellation (normal values scaled over 30 binades)."""
    import math
    x = orc.gen_normal(seed, n) * np.exp2(orc.gen_uniform_i32(seed + 1, -15, 15, n).astype(np.float64))
    hi, lo, ab = orc.prefix_sum_exact(x)
    rng = np.random.default_rng(seed)
    for i in sorted(rng.choice(n, 12, replace=False)) + [n - 1]:
        exact = math.fsum(x[: i + 1])
        assert abs((hi[i] + lo[i]) - exact) <= 2.0 ** -60 * ab[i] + 2.0 ** -52 * abs(exact), i


# ---------------------------------------------------------------- reduce
def test_reduce_known_answer_sequence_plus_13(orc):
    # reduce.cu:132-148: sequence(0..123456788) + 13 in int32 (wrapping)
    n = 123456789
    x = np.arange(n, dtype=np.int32)
    assert int(orc.reduce(x, 13)) == 1206807391
    closed = (n * (n - 1) // 2 + 13) % (1 << 32)
    assert (closed - (1 << 32) if closed >= (1 << 31) else closed) == 1206807391


def test_reduce_known_answer_sum_demo(orc):
    # sum.cu:63,69: 512 ones sum to 512
    assert int(orc.reduce(np.ones(512, np.int32), 0)) == 512


@pytest.mark.parametrize("op,fn", [("min", np.min), ("max", np.max)])
def test_reduce_min_max(orc, op, fn):
    x = orc.gen_normal(9, 3333)
    init = 0.0
    ref = fn(np.concatenate([[init], x]))
    assert orc.reduce(x, init, op) == ref


def test_sum_exact_matches_fsum(orc):
    import math
    x = orc.gen_normal(11, 10000)
    hi, lo, ab = orc.sum_exact(x)
    assert abs((hi + lo) - math.fsum(x)) <= 2.0 ** -60 * ab
    assert ab == pytest.approx(float(np.sum(np.abs(x))), rel=1e-12)


@pytest.mark.parametrize("seed,n,kind", [(12, 1, "u"), (13, 1000, "u"), (14, 100003, "u"),
                                         (15, 65536, "n"), (16, 300001, "n")])
def test_sum_exact_f32_matches_fsum(orc, seed, n, kind):
    """The float32 path (orc_sum_ld_f32, the checker of every f32 reduce):
    every float32 converts to float64 exactly, so math.fsum of the converted
    values is the exact sum rounded once; the long-double accumulator is
    within n * 2^-64 * sum|x| of it."""
    import math
    x = orc.gen_uniform_f32(seed, n) if kind == "u" else orc.gen_normal(seed, n, np.float32)
    assert x.dtype == np.float32
    hi, lo, ab = orc.sum_exact(x)
    exact = math.fsum(x.astype(np.float64))
    assert abs((hi + lo) - exact) <= n * 2.0 ** -63 * ab + 2.0 ** -52 * abs(exact)
    assert ab == math.fsum(np.abs(x).astype(np.float64))
    # hi is the sum rounded to double, lo the remainder
    assert hi == hi + lo or abs(lo) <= 2.0 ** -52 * abs(hi)


def test_sum_exact_f32_carries_more_than_double(orc):
    """2^60 + 1 - 2^60 is 1 exactly but 0 in float64 (and float32)
    accumulation: a checker that accumulated in double would miss the
    small terms a float32 reduce must keep."""
    x = np.array([2.0 ** 60, 1.0, -(2.0 ** 60)] * 5, dtype=np.float32)
    hi, lo, ab = orc.sum_exact(x)
    assert hi + lo == 5.0
    assert ab == 10 * 2.0 ** 60 + 5.0
    # float64 accumulation of the same values gives 0
    acc = 0.0
    for v in x.astype(np.float64):
        acc += v
    assert acc == 0.0


@pytest.mark.parametrize("n", [0, 1, 19999, 20000, 20001, 1152 * 37 + 5, 1 << 20])
@pytest.mark.parametrize("threads", [1, 2, 3, 8])
def test_mt_inclusive_scan_matches_single_thread(orc, n, threads):
    """orc_mt_inclusive_scan_i32 (the reference's 3-phase decomposition,
    scan.cu:109-145, one partition per thread) equals the one-thread body
    and a wrapping cumsum, bit for bit."""
    x = orc.gen_uniform_i32(7, -1000, 1000, n) if n else np.zeros(0, np.int32)
    got = orc.mt_inclusive_scan_i32(x, 13, threads)
    assert np.array_equal(got, orc.inclusive_scan(x, init=13))
    ref = (np.cumsum(x.astype(np.int64)) + 13).astype(np.int64)
    assert np.array_equal(got.astype(np.int64), ((ref + 2 ** 31) % 2 ** 32) - 2 ** 31)


def test_mt_inclusive_scan_wraps_like_int32(orc):
    x = np.full(100000, 2 ** 30, dtype=np.int32)
    got = orc.mt_inclusive_scan_i32(x, 0, 4)
    ref = ((np.arange(1, x.size + 1, dtype=np.int64) * 2 ** 30 + 2 ** 31) % 2 ** 32) - 2 ** 31
    assert np.array_equal(got.astype(np.int64), ref)


# ----------------------------------------------------------------- merge
@pytest.mark.parametrize("n1,n2", [(0, 0), (0, 5), (5, 0), (1, 1), (1000, 37), (4096, 4097)])
def test_merge_matches_stable_argsort(orc, n1, n2):
    a = np.sort(orc.ref_hash_i32(n1 + 3)[:n1] % 50)
    b = np.sort(orc.ref_hash_i32(n1 + n2 + 3)[n1:n1 + n2] % 50)
    out = orc.merge(a.astype(np.int32), b.astype(np.int32))
    cat = np.concatenate([a, b])
    assert np.array_equal(out, cat[np.argsort(cat, kind="stable")])


def test_merge_by_key_first_range_wins_ties(orc):
    pool = orc.gen_uniform_f32(2, 16)
    a = orc.gen_merge_keys(7, 0, 3000, pool, 0.5)
    b = orc.gen_merge_keys(7, 1, 2500, pool, 0.5)
    va = np.arange(a.size, dtype=np.int32)
    vb = (1 << 20) + np.arange(b.size, dtype=np.int32)
    ko, vo = orc.merge_by_key(a, va, b, vb)
    cat_k = np.concatenate([a, b])
    order = np.argsort(cat_k, kind="stable")
    assert np.array_equal(ko, cat_k[order])
    assert np.array_equal(vo, np.concatenate([va, vb])[order])


def test_merge_greater(orc):
    a = np.sort(orc.gen_uniform_i32(1, 0, 9, 100))[::-1].copy()
    b = np.sort(orc.gen_uniform_i32(2, 0, 9, 90))[::-1].copy()
    out = orc.merge(a, b, comp="greater")
    cat = np.concatenate([a, b])
    assert np.array_equal(out, cat[np.argsort(-cat.astype(np.int64), kind="stable")])


def test_merge_path_brute_force(orc):
    a = np.sort(orc.gen_uniform_i32(5, 0, 20, 60))
    b = np.sort(orc.gen_uniform_i32(6, 0, 20, 45))
    merged = orc.merge(a, b)
    src = np.argsort(np.concatenate([a, b]), kind="stable") < a.size
    for diag in range(a.size + b.size + 1):
        assert orc.merge_path(a, b, diag) == int(src[:diag].sum())
    assert merged.size == a.size + b.size


# ------------------------------------------------------------------ sort
@pytest.mark.parametrize("n", [0, 1, 6, 7, 8, 895, 896, 897, 5000, 20861])
def test_stable_sort_by_key_matches_stable_argsort(orc, n):
    keys = orc.ref_hash_i32(n)  # merge_sort_by_key.cu:264-286 input
    vals = np.arange(n, dtype=np.int32)
    k, v = orc.stable_sort_by_key(keys, vals)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(k, keys[order])
    assert np.array_equal(v, vals[order])


def test_stable_sort_few_distinct_keys_is_stable(orc):
    keys = (orc.gen_uniform_u32(3, 10000) % 7).astype(np.uint32)
    vals = np.arange(keys.size, dtype=np.uint32)
    k, v = orc.stable_sort_by_key(keys, vals, comp="greater")
    order = np.argsort(-keys.astype(np.int64), kind="stable")
    assert np.array_equal(k, keys[order]) and np.array_equal(v, vals[order])


# --------------------------------------------------------- reduce_by_key
@pytest.mark.parametrize("n,sms", [(1, 1), (2, 1), (19999, 1), (20000, 1), (20001, 1),
                                   (123457, 1), (123457, 3), (300000, 256)])
def test_reduce_by_key_interval_algorithm_matches_textbook(orc, n, sms):
    keys = orc.ref_hash_i32(n, mod10=True)  # reduce_by_key.cu:305-328 input
    vals = orc.ref_hash_i32(n, mod10=True)
    k1, v1 = orc.reduce_by_key(keys, vals, sms=sms)
    k2, v2 = orc.reduce_by_key(keys, vals, textbook=True)
    assert np.array_equal(k1, k2) and np.array_equal(v1, v2)


def test_reduce_by_key_textbook_brute_force(orc):
    keys = np.array([3, 3, 1, 1, 1, 2, 3, 3, 3, 0], np.int32)
    vals = np.array([1, 2, 3, 4, 5, 6, 7, 8, 9, 10], np.int64)
    k, v = orc.reduce_by_key(keys, vals, textbook=True)
    assert k.tolist() == [3, 1, 2, 3, 0] and v.tolist() == [3, 12, 6, 24, 10]


def test_reduce_by_key_long_runs_cross_intervals(orc):
    keys, runs = orc.gen_runs_zipf(4, 200000, 1.2, 1 << 16)
    vals = orc.gen_uniform_i64(4, -(1 << 20), 1 << 20, keys.size)
    k1, v1 = orc.reduce_by_key(keys, vals, sms=1)
    assert k1.size == runs
    bounds = np.flatnonzero(np.diff(keys)) + 1
    sums = np.add.reduceat(vals, np.concatenate([[0], bounds]))
    assert np.array_equal(v1, sums) and np.array_equal(k1, np.arange(runs, dtype=np.int32))


# ------------------------------------------------------ reduce_intervals
@pytest.mark.parametrize("n,P,tile", [(10, 3, 2), (1000, 7, 64), (5, 9, 1), (896 * 5 + 3, 4, 896)])
def test_aligned_decomposition_and_reduce_intervals(orc, n, P, tile):
    first, last = orc.aligned_ranges(n, P, tile)
    tiles = -(-n // tile)
    # partitions cover [0, n) in order, each a whole number of tiles except the clamp
    covered = np.concatenate([np.arange(f, l) for f, l in zip(first, last)])
    assert np.array_equal(covered, np.arange(n))
    counts = [(-(-(l - f) // tile)) for f, l in zip(first, last)]
    assert sorted(counts, reverse=True) == counts or max(counts) - min(counts) <= 1
    assert sum(counts) == tiles
    x = orc.gen_uniform_i32(8, -9, 9, n)
    out = orc.reduce_intervals(x, P, tile)
    for g in range(P):
        assert out[g] == (x[first[g]:last[g]].sum() if last[g] > first[g] else 0)


def test_flags_brute_force(orc):
    keys = np.array([1, 1, 2, 3, 3, 3, 1], np.int32)
    assert orc.tail_flags(keys).tolist() == [0, 1, 1, 0, 0, 1, 1]
    assert orc.head_flags(keys).tolist() == [1, 0, 1, 1, 0, 0, 1]


# ------------------------------------------------------------ generators
def test_generator_fixtures(orc):
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "generators.npz")
    g = np.load(path, allow_pickle=False)
    assert np.array_equal(orc.gen_uniform_i32(0, -1000, 1000, 64), g["uniform_i32_seed0"])
    assert np.array_equal(orc.gen_uniform_f32(0, 64), g["uniform_f32_seed0"])
    assert np.array_equal(orc.gen_normal(1, 64), g["normal_f64_seed1"])
    assert np.array_equal(orc.gen_runs_geometric(0, 256)[0], g["runs_geometric_seed0"])
    assert np.array_equal(orc.ref_hash_i32(64), g["ref_hash_i32"])
    assert np.array_equal(orc.ref_hash_i32(64, mod10=True), g["ref_hash_mod10_i32"])


def test_generator_distributions(orc):
    x = orc.gen_uniform_i32(0, -1000, 1000, 1 << 16)
    assert x.min() >= -1000 and x.max() <= 1000 and abs(x.mean()) < 10
    keys, runs = orc.gen_runs_geometric(0, 1 << 18, 8.0)
    assert abs((1 << 18) / runs - 8.0) < 0.2
    assert np.all(np.diff(keys) >= 0) and keys[-1] == runs - 1
    z, zr = orc.gen_runs_zipf(0, 1 << 20, 1.2, 1 << 16)
    assert np.all(np.diff(z) >= 0) and z[-1] == zr - 1


def test_oracle_signed_zero_keys(orc):
    """The oracle treats -0.0 and +0.0 as one key: its stable sort equals
    numpy's stable argsort (which compares with <), and its reduce_by_key
    writes each run's last key, bit for bit, as a plain loop does."""
    rng = np.random.default_rng(7)
    k = np.array([-1.0, -0.0, 0.0, 1.0], dtype=np.float32)[rng.integers(0, 4, 5000)]
    v = np.arange(k.size, dtype=np.int32)
    sk, sv = orc.stable_sort_by_key(k.copy(), v.copy())
    order = np.argsort(k, kind="stable")
    assert np.array_equal(sk.view(np.uint32), k[order].view(np.uint32))
    assert np.array_equal(sv, v[order])
    vals = rng.integers(-9, 9, k.size).astype(np.int32)
    ek, ev = orc.reduce_by_key(k, vals)
    tk, tv = [], []
    for i in range(k.size):
        if i > 0 and k[i] == k[i - 1]:
            tk[-1] = k[i]
            tv[-1] += int(vals[i])
        else:
            tk.append(k[i])
            tv.append(int(vals[i]))
    assert np.array_equal(ek.view(np.uint32), np.array(tk, dtype=np.float32).view(np.uint32))
    assert np.array_equal(ev, np.array(tv, dtype=np.int32))
