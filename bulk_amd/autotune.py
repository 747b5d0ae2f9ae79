"""Autotuner: measures every compiled tile shape of each algorithm on the
current GPU and writes the tuning table the library reads at start-up
(replacing the reference's hard-coded group/grain sizes and choose_sizes).

  python -m bulk_amd.autotune --out bulk_amd/tuning/mi355x.json

For each (algorithm, dtype, size class) the representative input (2^14,
2^18, 2^22, 2^24, 2^26 elements) is timed with HIP events under every variant
(median of `--iters` calls, workspace preallocated); the fastest variant is
stored.  reduce_by_key is timed on both run-length distributions of the
single-op benchmark (geometric mean 8 and Zipf), the two calls together.  Exclusive scan shares the inclusive scan's choice and merge (keys
only) the merge_by_key choice.  The batch executor's op families are tuned
on the benchmark's mixed-job stream (tune_batch): each family keeps the tile
shape under which its stage ran fastest.
"""
from __future__ import annotations

import argparse
import datetime
import json
import os
import sys

import torch

from . import _capi as C
from . import api
from . import opbench

SIZES = {0: 1 << 14, 1: 1 << 18, 2: 1 << 22, 3: 1 << 24, 4: 1 << 26}
DTYPES = {C.INT32: torch.int32, C.FLOAT32: torch.float32, C.INT64: torch.int64,
          C.FLOAT64: torch.float64}
WIDE_SCAN = 4  # the scan shape compiled for 4-byte integers only
DEFAULT_OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning", "mi355x.json")


def _time(fn, iters):
    for _ in range(2):
        fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(iters):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        torch.cuda.synchronize()
        ts.append(a.elapsed_time(b))
    ts.sort()
    return ts[len(ts) // 2]


def _variants(algo):
    c = C.ctypes.c_int()
    C.check(C.lib().bulkx_tuning_variants(algo, C.ctypes.byref(c)), "bulkx_tuning_variants")
    return c.value


def _set(algo, dtype, sc, v):
    C.check(C.lib().bulkx_set_tuning(algo, dtype, sc, v), "bulkx_set_tuning")


def _cases(g):
    """(algo, dtype code, size class, callable) for every measured case."""
    dev = "cuda"
    for sc, n in SIZES.items():
        for code in (C.INT32, C.FLOAT32, C.FLOAT64):
            dt = DTYPES[code]
            x = (torch.randint(-1000, 1001, (n,), generator=g, device=dev, dtype=dt)
                 if not dt.is_floating_point else torch.rand(n, generator=g, device=dev, dtype=dt))
            y = torch.empty_like(x)
            ws = api.Workspace.for_algo(C.ALGO_INCLUSIVE_SCAN, code, -1, n)
            yield C.ALGO_INCLUSIVE_SCAN, code, sc, (lambda x=x, y=y, ws=ws: api.inclusive_scan(x, y, workspace=ws))
            out = torch.empty(1, dtype=dt, device=dev)
            wsr = api.Workspace.for_algo(C.ALGO_REDUCE, code, -1, n)
            yield C.ALGO_REDUCE, code, sc, (lambda x=x, out=out, ws=wsr: api.reduce_into(x, out, 0, workspace=ws))
        for code in (C.INT32, C.FLOAT32, C.UINT32):
            dt = DTYPES.get(code, torch.int32)
            if code == C.UINT32:  # full-range unsigned sort keys (viewed bits)
                if not hasattr(torch, "uint32"):
                    continue
                k0 = torch.randint(-(1 << 31), (1 << 31) - 1, (n,), generator=g, device=dev,
                                   dtype=torch.int32).view(torch.uint32)
                k = k0.clone()
                v = torch.arange(n, dtype=torch.int32, device=dev)
                wss = api.Workspace.for_algo(C.ALGO_SORT_BY_KEY, code, C.INT32, n)

                def usrt(k=k, k0=k0, v=v, ws=wss):
                    k.copy_(k0)
                    api.stable_merge_sort_by_key(k, v, workspace=ws)
                yield C.ALGO_SORT_BY_KEY, code, sc, usrt
                continue
            h = n // 2
            a = torch.sort(torch.randint(-(1 << 30), 1 << 30, (h,), generator=g, device=dev).to(dt)).values
            b = torch.sort(torch.randint(-(1 << 30), 1 << 30, (n - h,), generator=g, device=dev).to(dt)).values
            va = torch.arange(h, dtype=torch.int32, device=dev)
            vb = torch.arange(n - h, dtype=torch.int32, device=dev)
            ko = torch.empty(n, dtype=dt, device=dev)
            vo = torch.empty(n, dtype=torch.int32, device=dev)
            yield C.ALGO_MERGE_BY_KEY, code, sc, (lambda a=a, va=va, b=b, vb=vb, ko=ko, vo=vo:
                                                  api.merge_by_key(a, va, b, vb, ko, vo))
            k0 = torch.randint(-(1 << 30), 1 << 30, (n,), generator=g, device=dev).to(dt)
            k = k0.clone()
            v = torch.arange(n, dtype=torch.int32, device=dev)
            wss = api.Workspace.for_algo(C.ALGO_SORT_BY_KEY, code, C.INT32, n)

            def srt(k=k, k0=k0, v=v, ws=wss):
                k.copy_(k0)
                api.stable_merge_sort_by_key(k, v, workspace=ws)
            yield C.ALGO_SORT_BY_KEY, code, sc, srt
        # the single-op benchmark's two run-length distributions, geometric
        # (mean 8: a run ends every few elements) and Zipf (long runs): each
        # variant is timed on both, one after the other
        keysets = [opbench._runs(n, kind, g, torch.device(dev)) for kind in ("geometric", "zipf")]
        for code in (C.INT32, C.FLOAT32, C.INT64):
            dt = DTYPES[code]
            vals = torch.ones(n, dtype=dt, device=dev)
            ko = torch.empty_like(keysets[0])
            vo = torch.empty_like(vals)
            cnt = torch.empty(1, dtype=torch.int64, device=dev)
            wsk = api.Workspace.for_algo(C.ALGO_REDUCE_BY_KEY, C.INT32, code, n)

            def rbk(keysets=keysets, vals=vals, ko=ko, vo=vo, cnt=cnt, ws=wsk):
                for keys in keysets:
                    api.reduce_by_key_async(keys, vals, ko, vo, cnt, workspace=ws)
            yield C.ALGO_REDUCE_BY_KEY, code, sc, rbk


def tune(iters: int = 10, log=print, only=None) -> dict:
    g = torch.Generator(device="cuda")
    g.manual_seed(0)
    entries = []
    for algo, code, sc, fn in _cases(g):
        if only is not None and algo not in only:
            continue
        times = []
        for v in range(_variants(algo)):
            if algo == C.ALGO_INCLUSIVE_SCAN and v == WIDE_SCAN and code not in (C.INT32, C.UINT32):
                times.append(float("inf"))  # runs variant 0's shape for this dtype
                continue
            _set(algo, code, sc, v)
            times.append(_time(fn, iters))
        best = min(range(len(times)), key=lambda i: times[i])
        _set(algo, code, sc, best)
        e = {"algo": algo, "dtype": code, "size_class": sc, "n": SIZES[sc], "variant": best,
             "ms": [round(t, 5) if t != float("inf") else None for t in times]}
        entries.append(e)
        log(f"algo {algo} dtype {code} n 2^{SIZES[sc].bit_length() - 1}: "
            f"{' '.join(f'{t * 1e3:.1f}us' for t in times)} -> variant {best}")
        shared = {C.ALGO_INCLUSIVE_SCAN: C.ALGO_EXCLUSIVE_SCAN, C.ALGO_MERGE_BY_KEY: C.ALGO_MERGE}
        if algo in shared:
            entries.append({**e, "algo": shared[algo], "copied_from": algo})
        torch.cuda.empty_cache()
    props = torch.cuda.get_device_properties(0)
    return {"device": props.name, "gcn_arch": getattr(props, "gcnArchName", ""),
            "generated": datetime.datetime.now(datetime.timezone.utc).isoformat(timespec="seconds"),
            "command": "python -m bulk_amd.autotune", "iters": iters, "entries": entries}


# The batch executor's families: stage name -> (tuning algo, dtype key).
BATCH_FAMILIES = {
    "scan_i32": (C.ALGO_BATCH_SCAN, C.INT32),
    "scan_f32": (C.ALGO_BATCH_SCAN, C.FLOAT32),
    "reduce_by_key_i32": (C.ALGO_BATCH_REDUCE_BY_KEY, C.INT32),
    "reduce_by_key_f32": (C.ALGO_BATCH_REDUCE_BY_KEY, C.FLOAT32),
    "reduce_by_key_i64": (C.ALGO_BATCH_REDUCE_BY_KEY, C.INT64),
    "merge_by_key": (C.ALGO_BATCH_MERGE, C.INT32),
}


def tune_batch(iters: int = 5, jobs: int = 4096, seed: int = 5, log=print) -> list[dict]:
    """The batch executor's family tile shapes, timed on the mixed-job
    stream the benchmark runs (`jobs` jobs, `seed`): every family's stage
    time (HIP events around its kernels in an eager launch, median of
    `iters`) under each shape; a family keeps its fastest shape."""
    from . import workload as W
    from .batch import Batch
    specs = W.make_specs(jobs, seed=seed)
    data = W.build_jobs(specs, seed=seed + 2)
    total = W.total_elements(data)
    nshapes = _variants(C.ALGO_BATCH_SCAN)
    times = {f: [] for f in BATCH_FAMILIES}
    for sh in range(nshapes):
        for algo, dt in BATCH_FAMILIES.values():
            _set(algo, dt, 0, sh)
        b = Batch(max_jobs=len(data), max_elements=total)
        b.reserve(total)
        b.set_jobs([j.job for j in data])
        runs = {f: [] for f in BATCH_FAMILIES}
        for _ in range(iters + 1):
            W.reset_all(data)
            torch.cuda.synchronize()
            b.launch(use_graph=False)
            st = b.stage_timing()
            for f in BATCH_FAMILIES:
                runs[f].append(st[f])
        assert b.status() == 0
        b.close()
        for f in BATCH_FAMILIES:
            r = sorted(runs[f][1:])
            times[f].append(r[len(r) // 2])
        log(f"batch shape {sh}: " + " ".join(f"{f} {times[f][-1]:.3f}ms" for f in BATCH_FAMILIES))
    entries = []
    for f, (algo, dt) in BATCH_FAMILIES.items():
        best = min(range(nshapes), key=lambda i: times[f][i])
        _set(algo, dt, 0, best)
        entries.append({"algo": algo, "dtype": dt, "size_class": 0, "family": f, "jobs": jobs,
                        "seed": seed, "variant": best, "ms": [round(t, 5) for t in times[f]]})
        log(f"batch family {f}: {' '.join(f'{t:.3f}ms' for t in times[f])} -> shape {best}")
    del data
    torch.cuda.empty_cache()
    return entries


def main(argv=None):
    p = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    p.add_argument("--out", default=DEFAULT_OUT)
    p.add_argument("--iters", type=int, default=10)
    p.add_argument("--no-batch", action="store_true", help="skip the batch executor's families")
    p.add_argument("--batch-only", action="store_true",
                   help="tune only the batch executor's families; keep the other entries of --out")
    p.add_argument("--only", default=None,
                   help="comma-separated algorithm codes to tune (bulkx_algo); the other entries "
                        "of --out are kept")
    a = p.parse_args(argv)
    C.lib()
    C.check(C.lib().bulkx_load_tuning(b"none", None), "bulkx_load_tuning")  # start from defaults
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    if a.batch_only:
        with open(a.out) as f:
            table = json.load(f)
        table["entries"] = [e for e in table["entries"] if e["algo"] < C.ALGO_BATCH_SCAN]
    elif a.only:
        only = {int(x) for x in a.only.split(",")}
        with open(a.out) as f:
            table = json.load(f)
        kept = [e for e in table["entries"] if e["algo"] not in only]
        fresh = tune(a.iters, log=log, only=only)
        table["entries"] = kept + fresh["entries"]
        table["generated"] = fresh["generated"]
        a.no_batch = True
    if not a.no_batch:
        table["entries"] += tune_batch(log=log)
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    with open(a.out, "w") as f:
        json.dump(table, f, indent=1)
    print(f"wrote {a.out} ({len(table['entries'])} entries)", flush=True)


if __name__ == "__main__":
    main()
