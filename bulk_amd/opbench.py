"""Single-op throughput at the full single-GPU sizes, each timed with HIP
events on the launching stream (median of `iters` calls, workspaces
preallocated), reported as GB/s of algorithmic bytes and as a fraction of
the 8 TB/s HBM peak:

  scan int32 2^24 uniform [-1000, 1000]                        2*n*4 bytes
  reduce / inclusive_scan 2^28 f32 and f64, uniform [0,1) and normal(0,1)
                                               reduce n*s, scan 2*n*s bytes
  merge_by_key 2 x 2^27 f32 keys (10% from a 1024-value pool) + i32 payload
                                                       2*(n1+n2)*8 bytes
  stable_sort_by_key 2^26 uint32 keys (full range) + uint32 index
                                          2*n*8 bytes (compulsory), passes
  reduce_by_key 2^28 int32 keys in runs (geometric mean 8; Zipf 1.2 capped
  at 2^16), values f32 uniform [-1,1] or int64 uniform [-2^20, 2^20]
                                         n*(4+s) + R*(4+s) + 8 bytes

Float results carry the measured error against a float64 reference
computed with torch on the GPU (float64 reduce: a double-double TwoSum tree),
as a fraction of the bound 4*ceil(log2 n)*eps*sum|x|.
"""
from __future__ import annotations

import math

import torch

from . import _capi as C
from . import api

HBM_PEAK_GBS = 8000.0


def hbm_ceiling() -> dict:
    """Run the HBM ceiling program (bulk_amd/lib/hbm_ceiling, built from
    bulk_amd/hbm_ceiling.hip) and return its JSON line: the GB/s of the best
    hand-written nontemporal read, copy and read-two-write-an-eighth kernels
    on 1 GiB streams, the practical ceilings of the kernels measured here."""
    import json
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "hbm_ceiling")
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} is not built (python -m bulk_amd.build)")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120, check=True)
    return json.loads(r.stdout.strip().splitlines()[-1])


def calibration_copy(dev, reps: int = 3) -> None:
    """Copies of 1 GiB (2^28 int32) in this process: a kernel of known
    traffic (1 GiB read, 1 GiB written) that scales the FETCH_SIZE and
    WRITE_SIZE counters of a rocprofv3 --pmc run of it."""
    src = torch.ones(1 << 28, dtype=torch.int32, device=dev)
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def _timeit(fn, iters=10, warm=2, stream=None):
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(iters):
        s = torch.cuda.Event(enable_timing=True)
        e = torch.cuda.Event(enable_timing=True)
        s.record()
        fn()
        e.record()
        torch.cuda.synchronize()
        ts.append(s.elapsed_time(e))
    ts.sort()
    return ts[len(ts) // 2]


def _timeit_back_to_back(fn, calls=20, iters=5):
    """Milliseconds per call of `calls` calls enqueued back to back between
    two events (median over `iters`): the device-side cost of a call once
    the host's launch latency is hidden behind the previous call."""
    fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(iters):
        s = torch.cuda.Event(enable_timing=True)
        e = torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(calls):
            fn()
        e.record()
        torch.cuda.synchronize()
        ts.append(s.elapsed_time(e) / calls)
    ts.sort()
    return ts[len(ts) // 2]


def _row(name, ms, nbytes, elements, **extra):
    gbs = nbytes / (ms * 1e-3) / 1e9
    return {"op": name, "ms": ms, "bytes": nbytes, "gbs": gbs, "frac_peak": gbs / HBM_PEAK_GBS,
            "melem_per_s": elements / (ms * 1e-3) / 1e6, **extra}


def _two_sum_tree(x: torch.Tensor) -> float:
    """float64 sum of x with a double-double (TwoSum) pairwise tree."""
    hi = x.clone()
    lo = torch.zeros_like(x)
    while hi.numel() > 1:
        if hi.numel() & 1:
            hi = torch.cat([hi, hi.new_zeros(1)])
            lo = torch.cat([lo, lo.new_zeros(1)])
        a, b = hi[0::2], hi[1::2]
        s = a + b
        bb = s - a
        err = (a - (s - bb)) + (b - bb)
        lo = lo[0::2] + lo[1::2] + err
        hi = s
    return float(hi[0]) + float(lo[0])


def _exact_sum(x: torch.Tensor) -> tuple[float, float]:
    """Sum of a float64 tensor as (hi, lo): a pairwise tree of TwoSum steps
    (Knuth) whose rounding errors are summed alongside; the error left is of
    order n * 2^-106 * sum|x|."""
    s = x
    e = torch.zeros_like(x)
    while s.numel() > 1:
        if s.numel() % 2:
            s = torch.cat([s, s.new_zeros(1)])
            e = torch.cat([e, e.new_zeros(1)])
        a, b = s[0::2], s[1::2]
        t = a + b
        bb = t - a
        err = (a - (t - bb)) + (b - bb)
        e = e[0::2] + e[1::2] + err
        s = t
    return float(s[0]), float(e[0])


def _f64_prefix_error(x64: torch.Tensor, y: torch.Tensor, eps: float, samples: int = 32) -> dict:
    """Error of a float64 scan at `samples` positions (and the last): the
    exact prefix is the sum of exactly summed segments between positions."""
    n = x64.numel()
    g = torch.Generator(device="cpu").manual_seed(n)
    pos = sorted(set(torch.randint(0, n, (samples,), generator=g).tolist()) | {n - 1})
    worst_err, worst_ratio, start, parts, absum = 0.0, 0.0, 0, [], 0.0
    for i in pos:
        seg = x64[start:i + 1]
        hi, lo = _exact_sum(seg)
        parts += [hi, lo]
        absum += float(seg.abs().sum())
        start = i + 1
        exact = math.fsum(parts)
        err = abs(float(y[i]) - exact)
        worst_err = max(worst_err, err)
        worst_ratio = max(worst_ratio, err / _bound(i + 1, eps, absum))
    return {"err": worst_err, "err_over_bound": worst_ratio, "err_sampled_positions": len(pos)}


def _sort_ok(k: torch.Tensor, v: torch.Tensor, k0: torch.Tensor) -> bool:
    """Sorted output of a stable sort by key with payload = input index:
    keys nondecreasing, payload a permutation with key == input key at it,
    equal keys in input order."""
    ki = k.view(torch.int32).long() & 0xFFFFFFFF
    vi = v.view(torch.int32).long()
    n = ki.numel()
    seen = torch.zeros(n, dtype=torch.uint8, device=k.device)
    seen[vi] = 1
    ties = ki[1:] == ki[:-1]
    return bool((ki[1:] >= ki[:-1]).all()) and bool(seen.all()) and \
        bool(torch.equal(k.view(torch.int32), k0.view(torch.int32)[vi])) and \
        bool((vi[1:][ties] > vi[:-1][ties]).all())


def _bound(n, eps, absum):
    return 4 * max(1, math.ceil(math.log2(max(n, 2)))) * eps * absum


def _zipf_lengths(m, s, cap, g, dev):
    k = torch.arange(1, cap + 1, dtype=torch.float64, device=dev)
    cdf = torch.cumsum(k.pow(-s), 0)
    cdf /= cdf[-1].clone()
    u = torch.rand(m, generator=g, device=dev, dtype=torch.float64)
    return (torch.searchsorted(cdf, u) + 1).clamp_(max=cap)


def _runs(n, kind, g, dev):
    if kind == "geometric":
        m = int(n / 8 * 1.2) + 64
        lens = torch.empty(m, device=dev).geometric_(1 / 8, generator=g).to(torch.int64)
    else:
        m = int(n / 1766 * 1.5) + 64
        lens = _zipf_lengths(m, 1.2, 1 << 16, g, dev)
    assert int(lens.sum()) >= n
    keys = torch.repeat_interleave(torch.arange(m, dtype=torch.int32, device=dev), lens)[:n]
    return keys.contiguous()


def run(iters: int = 10, log=print, large: int = 28) -> list[dict]:
    dev = torch.device("cuda")
    g = torch.Generator(device=dev)
    rows = []
    ceil = hbm_ceiling()
    calibration_copy(dev)
    log(f"HBM ceilings: read {ceil['read_nt_gbs']:.0f} GB/s, copy {ceil['copy_nt_gbs']:.0f} GB/s "
        f"(plain stores {ceil['copy_plain_gbs']:.0f}), read two + write an eighth "
        f"{ceil['read2_write8_gbs']:.0f} GB/s")

    # scan int32 2^24
    g.manual_seed(0)
    n = 1 << 24
    x = torch.randint(-1000, 1001, (n,), generator=g, device=dev, dtype=torch.int32)
    y = torch.empty_like(x)
    ws = api.Workspace.for_algo(C.ALGO_INCLUSIVE_SCAN, C.INT32, -1, n)
    call = lambda: api.inclusive_scan(x, y, workspace=ws)  # noqa: E731
    rows.append(_row("inclusive_scan int32 2^24", _timeit(call, iters), 2 * n * 4, n,
                     exact=bool(torch.equal(y, torch.cumsum(x, 0, dtype=torch.int32))),
                     ms_back_to_back=_timeit_back_to_back(call, 20, iters)))
    del x, y, ws

    # reduce / scan at 2^large in f32 / f64, uniform and normal
    n = 1 << large
    for dt, eps in ((torch.float32, 2.0 ** -24), (torch.float64, 2.0 ** -53)):
        for dist in ("uniform", "normal"):
            g.manual_seed(0 if dist == "uniform" else 1)
            x = (torch.rand(n, generator=g, device=dev, dtype=torch.float64) if dist == "uniform"
                 else torch.randn(n, generator=g, device=dev, dtype=torch.float64)).to(dt)
            tag = f"{'f32' if dt == torch.float32 else 'f64'} 2^{large} {dist}"
            out = torch.empty(1, dtype=dt, device=dev)
            wsr = api.Workspace.for_algo(C.ALGO_REDUCE, api.dtype_code(x), -1, n)
            ms = _timeit(lambda: api.reduce_into(x, out, 0, workspace=wsr), iters)
            x64 = x.double()
            absum = float(x64.abs().sum())
            ref = _two_sum_tree(x64)
            err = abs(float(out.item()) - ref)
            rows.append(_row(f"reduce {tag}", ms, n * x.element_size(), n,
                             err=err, err_over_bound=err / _bound(n, eps, absum)))
            del wsr
            y = torch.empty_like(x)
            wss = api.Workspace.for_algo(C.ALGO_INCLUSIVE_SCAN, api.dtype_code(x), -1, n)
            ms = _timeit(lambda: api.inclusive_scan(x, y, workspace=wss), iters)
            extra = {}
            if dt == torch.float32:  # float64 cumsum is an adequate reference for float32
                p = torch.cumsum(x64, 0)
                a = torch.cumsum(x64.abs(), 0)
                e = (y.double() - p).abs()
                b = 4 * max(1, math.ceil(math.log2(n))) * eps * a
                extra = {"err": float(e.max()), "err_over_bound": float((e / b.clamp_min(1e-300)).max())}
                del p, a, e, b
            else:  # float64: exact prefixes at sampled positions (a float64 cumsum is too coarse)
                extra = _f64_prefix_error(x64, y, eps)
            rows.append(_row(f"inclusive_scan {tag}", ms, 2 * n * x.element_size(), n, **extra))
            del x, y, x64, wss, out
            torch.cuda.empty_cache()

    # merge_by_key 2 x 2^(large-1) f32 + i32 payload
    h = 1 << (large - 1)
    g.manual_seed(2)
    pool = torch.rand(1024, generator=g, device=dev)

    def mkeys():
        k = torch.rand(h, generator=g, device=dev)
        sel = torch.rand(h, generator=g, device=dev) < 0.1
        k = torch.where(sel, pool[torch.randint(0, 1024, (h,), generator=g, device=dev)], k)
        return torch.sort(k).values

    a, b = mkeys(), mkeys()
    va = torch.arange(h, dtype=torch.int32, device=dev)
    vb = va + h
    ko = torch.empty(2 * h, device=dev)
    vo = torch.empty(2 * h, dtype=torch.int32, device=dev)
    wsm = api.Workspace.for_algo(C.ALGO_MERGE_BY_KEY, C.FLOAT32, C.INT32, h, h)
    ms = _timeit(lambda: api.merge_by_key(a, va, b, vb, ko, vo, workspace=wsm), iters)
    ok = bool(torch.equal(ko, torch.sort(torch.cat([a, b]), stable=True).values))
    rows.append(_row(f"merge_by_key f32+i32 2 x 2^{large - 1}", ms, 2 * (2 * h) * 8, 2 * h,
                     keys_sorted=ok))
    del a, b, va, vb, ko, vo, pool
    torch.cuda.empty_cache()

    # stable_sort_by_key 2^(large-2) uint32 + index
    s = 1 << (large - 2)
    g.manual_seed(3)
    k0 = torch.randint(-(1 << 31), (1 << 31) - 1, (s,), generator=g, device=dev, dtype=torch.int32)
    k0 = k0.view(torch.uint32) if hasattr(torch, "uint32") else k0
    v0 = torch.arange(s, dtype=torch.int32, device=dev).view(k0.dtype)
    k, v = k0.clone(), v0.clone()
    wst = api.Workspace.for_algo(C.ALGO_SORT_BY_KEY, api.dtype_code(k), api.dtype_code(v), s)
    t_copy = _timeit(lambda: (k.copy_(k0), v.copy_(v0)), iters)

    def srt():
        k.copy_(k0)
        v.copy_(v0)
        api.stable_merge_sort_by_key(k, v, workspace=wst)

    ms = _timeit(srt, iters) - t_copy
    variant = api.tuned_variant(C.ALGO_SORT_BY_KEY, api.dtype_code(k), s)
    if variant == 4:  # radix: 4 passes, each an upsweep (keys) and a downsweep (keys + values)
        form = {"form": "radix", "passes": 4, "pass_bytes": 4 * (4 * s + 2 * s * 8)}
    else:
        _, gs, gr = api.choose_sizes(C.ALGO_SORT_BY_KEY, api.dtype_code(k), s)  # block-sorted run
        runs = (s + gs * gr - 1) // (gs * gr)
        passes = max(0, math.ceil(math.log2(runs)))
        form = {"form": "merge", "passes": passes, "pass_bytes": (1 + passes) * 2 * s * 8}
    rows.append(_row(f"stable_sort_by_key u32+u32 2^{large - 2}", ms, 2 * s * 8, s,
                     sorted_stable_permutation=_sort_ok(k, v, k0), **form,
                     note="time excludes restoring the unsorted input (timed separately)"))
    del k0, v0, k, v, wst
    torch.cuda.empty_cache()

    # reduce_by_key 2^large int32 keys
    n = 1 << large
    for runs in ("geometric", "zipf"):
        g.manual_seed(4)
        keys = _runs(n, runs, g, dev)
        R = int(keys[-1]) + 1
        for vdt in (torch.float32, torch.int64):
            if vdt == torch.float32:
                vals = torch.rand(n, generator=g, device=dev) * 2 - 1
            else:
                vals = torch.randint(-(1 << 20), (1 << 20) + 1, (n,), generator=g, device=dev,
                                     dtype=torch.int64)
            ko = torch.empty_like(keys)
            vo = torch.empty_like(vals)
            cnt = torch.empty(1, dtype=torch.int64, device=dev)
            wsk = api.Workspace.for_algo(C.ALGO_REDUCE_BY_KEY, C.INT32, api.dtype_code(vals), n)
            ms = _timeit(lambda: api.reduce_by_key_async(keys, vals, ko, vo, cnt, workspace=wsk), iters)
            es = vals.element_size()
            ok = int(cnt.item()) == R and bool(torch.equal(ko[:R], torch.arange(R, dtype=torch.int32,
                                                                                 device=dev)))
            extra = {"runs": R, "keys_and_count_exact": ok}
            if vdt == torch.int64:
                ref = torch.zeros(R, dtype=torch.int64, device=dev).index_add_(0, keys.long(), vals)
                extra["sums_exact"] = bool(torch.equal(vo[:R], ref))
            else:  # float64 index_add errs by < 2^-37 sum|x| per run: far below the bound
                idx = keys.long()
                v64 = vals.double()
                ref = torch.zeros(R, dtype=torch.float64, device=dev).index_add_(0, idx, v64)
                absum = torch.zeros(R, dtype=torch.float64, device=dev).index_add_(0, idx, v64.abs())
                lens = torch.bincount(idx, minlength=R).double()
                bnd = 4 * torch.clamp(torch.ceil(torch.log2(torch.clamp(lens, min=2))), min=1) \
                    * 2.0 ** -24 * absum
                err = (vo[:R].double() - ref).abs()
                extra["err"] = float(err.max())
                extra["err_over_bound"] = float((err / bnd.clamp_min(1e-300)).max())
                del idx, v64, ref, absum, lens, bnd, err
            rows.append(_row(f"reduce_by_key i32/{'f32' if vdt == torch.float32 else 'i64'} 2^{large} {runs}",
                             ms, n * (4 + es) + R * (4 + es) + 8, n, **extra))
            del vals, ko, vo, wsk
        del keys
        torch.cuda.empty_cache()
    for r in rows:
        r["frac_of_copy_ceiling"] = r["gbs"] / ceil["copy_nt_gbs"]
    rows.insert(0, {"op": "hbm_ceiling", **ceil})
    for r in rows[1:]:
        b2b = f"  (back to back {r['ms_back_to_back'] * 1e3:.1f} us)" if "ms_back_to_back" in r else ""
        log(f"{r['op']:46s} {r['ms'] * 1e3:9.1f} us {r['gbs']:7.0f} GB/s {100 * r['frac_peak']:5.1f}% "
            f"({100 * r['frac_of_copy_ceiling']:5.1f}% of copy){b2b}")
    return rows


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser(description="single-op table at full size (one MI355X)")
    ap.add_argument("--iters", type=int, default=10)
    ap.add_argument("--large", type=int, default=28)
    ap.add_argument("--json", default=None, help="also write the rows to this file")
    a = ap.parse_args()
    rows = run(iters=a.iters, large=a.large)
    if a.json:
        import json

        with open(a.json, "w") as f:
            json.dump(rows, f, indent=1)
