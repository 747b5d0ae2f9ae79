"""GPU parity: inclusive/exclusive scan, reduce and reduce_intervals through
the C-ABI against the CPU oracle.  Integer results are bit-exact; float
results are within 4*ceil(log2 n)*eps*sum|x| of the extended-precision
oracle (eps = 2^-24 for float32, 2^-53 for float64)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

EPS = {np.float32: 2.0 ** -24, np.float64: 2.0 ** -53}
EDGE_SIZES = [1, 2, 63, 64, 65, 255, 4095, 4096, 4097, 8191, 19999, 20000, 20001, 65543]
MINSTD_SIZES = [48271, 153570, 597830, 20861]


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _host(t):
    return t.cpu().numpy()


def _ints(orc, dt, n, seed):
    if dt == np.int32:
        return orc.gen_uniform_i32(seed, -1000, 1000, n)
    if dt == np.uint32:
        return orc.gen_uniform_u32(seed, n)
    if dt == np.int64:
        return orc.gen_uniform_i64(seed, -(1 << 40), 1 << 40, n)
    return orc.gen_uniform_u32(seed, n).astype(np.uint64) << np.uint64(20)


@pytest.mark.parametrize("n", EDGE_SIZES + MINSTD_SIZES)
@pytest.mark.parametrize("dt", [np.int32, np.int64])
def test_inclusive_scan_integer_bit_exact(bx, orc, n, dt):
    print(f"scan int dt={np.dtype(dt).name} n={n}", flush=True)
    x = _ints(orc, dt, n, seed=n)
    out = _host(bx.inclusive_scan(_dev(x), init=13))
    assert np.array_equal(out, orc.inclusive_scan(x, init=13))
    out = _host(bx.inclusive_scan(_dev(x)))
    assert np.array_equal(out, orc.inclusive_scan(x))


@pytest.mark.parametrize("dt", [np.uint32, np.uint64])
def test_inclusive_scan_unsigned_wraps(bx, orc, dt):
    if not hasattr(torch, "uint32"):
        pytest.skip("torch lacks unsigned dtypes")
    x = _ints(orc, dt, 100003, seed=3)
    out = _host(bx.inclusive_scan(_dev(x), init=5))
    assert np.array_equal(out, orc.inclusive_scan(x, init=5))


@pytest.mark.parametrize("op", ["min", "max"])
@pytest.mark.parametrize("dt", [np.int32, np.int64, np.float32, np.float64])
def test_inclusive_scan_min_max_exact(bx, orc, op, dt):
    n = 300007
    x = (orc.gen_normal(7, n) * 1000).astype(dt)
    out = _host(bx.inclusive_scan(_dev(x), op=op))
    assert np.array_equal(out, orc.inclusive_scan(x, op=op))


@pytest.mark.parametrize("n", [1, 4097, 1 << 20, 20861])
def test_exclusive_scan_bit_exact(bx, orc, n):
    x = _ints(orc, np.int32, n, seed=11)
    out = _host(bx.exclusive_scan(_dev(x), init=13))
    assert np.array_equal(out, orc.exclusive_scan(x, init=13))


def test_scan_known_answer_ones_plus_13(bx):
    # scan.cu:158-184
    for n in [1 << k for k in range(0, 21, 4)] + [123457]:
        x = torch.ones(n, dtype=torch.int32, device="cuda")
        out = bx.inclusive_scan(x, init=13)
        assert torch.equal(out, 14 + torch.arange(n, dtype=torch.int32, device="cuda"))


def test_scan_in_place(bx, orc):
    x = _ints(orc, np.int32, 1 << 18, seed=2)
    d = _dev(x)
    bx.inclusive_scan(d, d, init=-7)
    assert np.array_equal(_host(d), orc.inclusive_scan(x, init=-7))


def test_scan_config1_int32_2p24_seed0(bx, orc):
    # int32 uniform [-1000, 1000], seed 0, n = 2^24: bit-exact
    x = orc.gen_uniform_i32(0, -1000, 1000, 1 << 24)
    ref = orc.inclusive_scan(x, init=0)
    out = _host(bx.inclusive_scan(_dev(x), init=0))
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("dist", ["uniform", "normal"])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("n", [1000, 4097, 1 << 22])
def test_scan_float_within_bound(bx, orc, dt, dist, n):
    print(f"scan float {np.dtype(dt).name} {dist} n={n}", flush=True)
    x = (orc.gen_uniform_f64(0, n) if dist == "uniform" else orc.gen_normal(1, n)).astype(dt)
    hi, lo, ab = orc.prefix_sum_exact(x)
    out = _host(bx.inclusive_scan(_dev(x))).astype(np.float64)
    err = np.abs((out - hi) - lo)
    bound = orc.float_tolerance(n, EPS[dt], ab)
    assert np.all(err <= bound), f"max err/bound {np.max(err / np.maximum(bound, 1e-300))}"


@pytest.mark.parametrize("n", EDGE_SIZES + [1 << 22, 123456789 // 16])
@pytest.mark.parametrize("dt", [np.int32, np.int64])
def test_reduce_integer_bit_exact(bx, orc, n, dt):
    x = _ints(orc, dt, n, seed=n + 1)
    assert bx.my_reduce(_dev(x), 13) == int(orc.reduce(x, 13))


def test_reduce_known_answer_sequence_plus_13(bx):
    # reduce.cu:132-148
    x = torch.arange(123456789, dtype=torch.int32, device="cuda")
    assert bx.my_reduce(x, 13) == 1206807391


def test_reduce_unaligned_views(bx, orc):
    x = _ints(orc, np.int32, 100000, seed=5)
    d = _dev(x)
    for off in range(1, 5):
        view = d[off:]
        out = torch.empty(1, dtype=torch.int32, device="cuda")
        bx.reduce_into(view, out, 0)
        assert int(out.item()) == int(orc.reduce(x[off:], 0))


@pytest.mark.parametrize("op", ["min", "max"])
@pytest.mark.parametrize("dt", [np.int32, np.float32, np.float64, np.int64])
def test_reduce_min_max_exact(bx, orc, op, dt):
    x = (orc.gen_normal(3, 1000003) * 1e6).astype(dt)
    init = dt(0)
    assert bx.my_reduce(_dev(x), init, op) == orc.reduce(x, init, op)


@pytest.mark.parametrize("dist", ["uniform", "normal"])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("n", [1, 1000, 1 << 20, (1 << 24) + 3])
def test_reduce_float_within_bound(bx, orc, dt, dist, n):
    x = (orc.gen_uniform_f64(2, n) if dist == "uniform" else orc.gen_normal(4, n)).astype(dt)
    hi, lo, ab = orc.sum_exact(x)
    got = float(bx.my_reduce(_dev(x), 0))
    err = abs((got - hi) - lo)
    assert err <= float(orc.float_tolerance(n, EPS[dt], ab))


def test_reduce_empty_returns_init(bx):
    x = torch.empty(0, dtype=torch.float32, device="cuda")
    assert bx.my_reduce(x, 13.5) == 13.5


@pytest.mark.parametrize("n,P,tile", [(10, 3, 2), (100000, 7, 640), (1 << 20, 2560, 896),
                                      (5, 9, 1), (123457, 13422, 640)])
def test_reduce_intervals(bx, orc, n, P, tile):
    x = _ints(orc, np.int32, n, seed=P)
    out = _host(bx.reduce_intervals(_dev(x), P, tile))
    assert np.array_equal(out[:P], orc.reduce_intervals(x, P, tile))
