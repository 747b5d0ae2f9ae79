"""GPU parity of the batch executor's sort jobs against a stable argsort on
the host, bit for bit (keys and payload = original index), on inputs beyond
the mixed stream's: uniform int32 / uint32 keys, uniform and normal float32
keys, keys with few distinct values, all-equal and heavily skewed keys, float
specials (NaN of both signs, +-inf, +-0, subnormals), a narrow float range,
presorted and reversed inputs, at sizes on both sides of the one-tile block
sort."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _order_image(keys: np.ndarray) -> np.ndarray:
    """The sort order of the library's 4-byte keys as an unsigned image:
    ints with the sign bit flipped, floats sign-magnitude to two's complement
    with -0.0 folded onto +0.0 (NaNs by their bits: sign set before -inf,
    clear after +inf)."""
    if keys.dtype == np.uint32:
        return keys.copy()
    if keys.dtype == np.int32:
        return keys.view(np.uint32) ^ np.uint32(0x80000000)
    b = keys.view(np.uint32).copy()
    b[b == 0x80000000] = 0
    neg = (b & np.uint32(0x80000000)) != 0
    return np.where(neg, ~b, b | np.uint32(0x80000000)).astype(np.uint32)


def _cases(rng):
    out = []
    i32 = lambda n: rng.integers(-(1 << 31), 1 << 31, size=n, dtype=np.int64).astype(np.int32)  # noqa: E731
    out.append(("i32 8193", i32(8193)))
    out.append(("i32 200001", i32(200001)))
    out.append(("i32 2^20", i32(1 << 20)))
    out.append(("u32 123457", rng.integers(0, 1 << 32, size=123457, dtype=np.uint64).astype(np.uint32)))
    out.append(("f32 unif 300000", rng.random(300000, dtype=np.float32)))
    out.append(("f32 normal 2^20", rng.standard_normal(1 << 20).astype(np.float32)))
    out.append(("i32 16 values 200000", rng.integers(0, 16, size=200000).astype(np.int32)))
    out.append(("i32 all equal 100000", np.full(100000, 7, np.int32)))
    sk = np.full(200000, 5, np.int32)
    sel = rng.random(200000) < 0.1
    sk[sel] = i32(int(sel.sum()))
    out.append(("i32 90% one value", sk))
    sp = rng.standard_normal(150000).astype(np.float32)
    specials = np.array([np.nan, -np.nan, np.inf, -np.inf, 0.0, -0.0, 1e-45, -1e-45, 1e-40,
                         3.4e38, -3.4e38], dtype=np.float32)
    idx = rng.integers(0, sp.size, size=3000)
    sp[idx] = specials[rng.integers(0, specials.size, size=idx.size)]
    nb = sp.view(np.uint32)
    nb[idx[:50]] = 0x7FC00001  # positive NaN payloads
    nb[idx[50:100]] = 0xFFC00002  # negative NaN payloads
    out.append(("f32 specials 150000", sp))
    out.append(("f32 narrow range 50000",
                (np.float32(1.0) + rng.random(50000, dtype=np.float32) * np.float32(1e-6)).astype(np.float32)))
    out.append(("i32 sorted 90000", np.sort(i32(90000))))
    out.append(("i32 reversed 90000", np.sort(i32(90000))[::-1].copy()))
    out.append(("f32 specials 5000 (one tile)", sp[:5000].copy()))
    out.append(("i32 all equal 8000 (one tile)", np.full(8000, -3, np.int32)))
    return out


@pytest.mark.parametrize("mode", ["graph", "eager"])
def test_batch_sort_matches_stable_argsort(bx, mode):
    from bulk_amd import batch as B
    rng = np.random.default_rng(11)
    cases = _cases(rng)
    jobs, refs = [], []
    for name, keys in cases:
        vals = np.arange(keys.size, dtype=np.uint32)
        if keys.dtype == np.uint32:
            dk = _dev(keys.view(np.int32)).view(torch.uint32)
        else:
            dk = _dev(keys)
        dv = _dev(vals.view(np.int32))
        order = np.argsort(_order_image(keys), kind="stable")
        jobs.append(B.sort_by_key_job(dk, dv))
        refs.append((name, dk, dv, keys.view(np.uint32)[order], vals[order]))
    total = sum(k.size for _, k in cases)
    b = B.Batch(max_jobs=len(jobs), max_elements=total)
    b.set_jobs(jobs)
    print(f"batch sort mode={mode} jobs={len(jobs)}", flush=True)
    b.launch(1 if mode == "graph" else 0)
    torch.cuda.synchronize()
    assert b.status() == 0
    bad = []
    for name, dk, dv, rk, rv in refs:
        got_k = dk.view(torch.int32).cpu().numpy().view(np.uint32)
        got_v = dv.cpu().numpy().view(np.uint32)
        if not (np.array_equal(got_k, rk) and np.array_equal(got_v, rv)):
            bad.append(name)
    b.close()
    assert bad == [], bad
