"""Mixed-job stream benchmark on one MI355X.

Workload: 4096 device-resident jobs (scan 40%, reduce_by_key 30%, merge 20%,
sort_by_key 10%), sizes log-uniform on [2^10, 2^22] elements, int32 or
float32 with probability 1/2 (bulk_amd.workload).

One step = the whole stream through the batch executor: one hipGraph replay
of the persistent segmented launch (clear + phase-0 + sort merge rounds).
The same job list is also run as one plain launch per job (the per-call
baseline) and through the serving loop (continuous batching under a 32 GB
HBM budget, saturated and with Poisson arrivals) for p50/p99 latency.

Prints ONE JSON line (rank 0, N=1).  `--gpus N` with N > 1 exits: this
build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HBM_PEAK_GBS = 8000.0


def _parse():
    p = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--jobs", type=int, default=4096)
    p.add_argument("--seed", type=int, default=5)
    p.add_argument("--budget-gb", type=float, default=32.0)
    p.add_argument("--no-serve", action="store_true",
                   help="skip the serving runs (batched and per-call steps only)")
    p.add_argument("--load", type=float, default=0.5,
                   help="Poisson arrival rate as a fraction of the saturated served job rate")
    p.add_argument("--cpu-seconds", type=float, default=10.0,
                   help="wall-clock budget of the CPU baseline sample")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-ops", action="store_true",
                   help="skip the single-op table (scan/reduce/merge/sort/reduce_by_key at full size)")
    p.add_argument("--no-per-call", action="store_true",
                   help="skip the per-call leg (and the batched == per-call check)")
    p.add_argument("--traffic-file", default=os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                          "profiles", "mixed_stream_4096_batch_traffic.md"))
    return p.parse_args()


def _log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _cpu_model():
    """The host CPU's model name and the cores this process may run on."""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": name, "visible_cores": len(os.sched_getaffinity(0)), "logical_cpus": os.cpu_count()}


def _pct(xs, q):
    import numpy as np
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q)) if len(xs) else None



def _read_traffic(path, kernel):
    """HBM bytes per launch of `kernel` from a committed rocprofv3 --pmc
    summary: the lines after `kernel: <name>` give `fetch_bytes_per_launch: X`
    and `write_bytes_per_launch: Y` (counter KiB scaled by a copy of known
    size), plus the `jobs` and `seed` of the workload they were measured on."""
    if not os.path.exists(path):
        return None
    vals, current = {}, None
    for line in open(path):
        if line.startswith("kernel:"):
            current = line.split(":", 1)[1].strip()
            continue
        if ":" in line and (current is None or current == kernel):
            k, v = line.split(":", 1)
            try:
                vals[k.strip()] = float(v.split()[0])
            except (ValueError, IndexError):
                pass
    if "fetch_bytes_per_launch" in vals and "write_bytes_per_launch" in vals:
        return vals["fetch_bytes_per_launch"] + vals["write_bytes_per_launch"], vals.get("jobs"), \
            vals.get("seed")
    return None


def main():
    args = _parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(jobs would shard by job across GPUs with no exchange).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from bulk_amd import _capi as C
    from bulk_amd import workload as W
    from bulk_amd.batch import Batch

    C.lib()  # fails loudly when the native library is not built
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)

    # ---------------------------------------------------------------- data
    t0 = time.perf_counter()
    specs = W.make_specs(args.jobs, seed=args.seed)
    jobs = W.build_jobs(specs, seed=args.seed + 2)
    torch.cuda.synchronize()
    total = W.total_elements(jobs)
    njobs = len(jobs)
    _log(f"generated {len(jobs)} jobs, {total} elements in {time.perf_counter() - t0:.1f}s")

    # HBM ceilings: the hand-written nontemporal copy and read kernels of
    # bulk_amd/hbm_ceiling.hip on 1 GiB streams (run as a separate process)
    from bulk_amd.opbench import calibration_copy, hbm_ceiling
    ceiling = hbm_ceiling()
    calibration_copy(dev)  # the traffic counters' scale under rocprofv3 --pmc
    copy_gbs = ceiling["copy_nt_gbs"]
    _log(f"device copy {copy_gbs:.0f} GB/s, read {ceiling['read_nt_gbs']:.0f} GB/s")

    # ----------------------------------------------------- per-call baseline
    def timed(fn, steps):
        ms = []
        for _ in range(steps):
            W.reset_all(jobs)
            torch.cuda.synchronize()
            a = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ms.append((time.perf_counter() - a) * 1e3)
        return ms

    ref_out, pc_ms = None, None
    if not args.no_per_call:
        pc = Batch(max_jobs=len(jobs), max_elements=total)
        pc.set_jobs([j.job for j in jobs])
        W.reset_all(jobs)
        pc.launch_per_call()
        torch.cuda.synchronize()
        ref_out = {}
        for i, j in enumerate(jobs):
            t = j.job.tensors
            if j.spec.kind == C.JOB_REDUCE_BY_KEY:
                m = int(t["out2"].item())
                ref_out[i] = (t["out0"][:m].clone(), t["out1"][:m].clone(), m)
            elif j.spec.kind == C.JOB_SORT_BY_KEY:
                ref_out[i] = (t["out0"].clone(), t["out1"].clone())
            else:
                ref_out[i] = tuple(t[k].clone() for k in ("out0", "out1") if k in t)
        for _ in range(args.warmup):
            W.reset_all(jobs)
            pc.launch_per_call()
        torch.cuda.synchronize()
        pc_ms = timed(pc.launch_per_call, args.steps)
        pc.close()
        _log(f"per-call: {np.mean(pc_ms):.2f} ms/step")

    # --------------------------------------------------------- batched (graph)
    bt = Batch(max_jobs=len(jobs), max_elements=total)
    bt.set_jobs([j.job for j in jobs])
    for _ in range(args.warmup):
        W.reset_all(jobs)
        bt.launch(use_graph=True)
    torch.cuda.synchronize()

    # equal accuracy: every batched result equals the per-call one (integer
    # work bit-exact, float scans / sums within the stated bound of FP64)
    W.reset_all(jobs)
    bt.launch(use_graph=True)
    torch.cuda.synchronize()
    runs = {i: int(j.job.tensors["out2"].item()) for i, j in enumerate(jobs)
            if j.spec.kind == C.JOB_REDUCE_BY_KEY}
    mismatches, worst_ratio = 0, 0.0
    for i, j in enumerate(jobs):
        if ref_out is None:
            break
        t = j.job.tensors
        k, dt, n = j.spec.kind, j.spec.dtype, j.spec.n
        if k == C.JOB_REDUCE_BY_KEY:
            m = runs[i]
            ok = m == ref_out[i][2] and torch.equal(t["out0"][:m], ref_out[i][0])
            if dt == "int32":
                ok = ok and torch.equal(t["out1"][:m], ref_out[i][1])
            else:  # FP64 run sums (keys are run indices 0..m-1)
                kk = t["in0"].long()
                v64 = torch.zeros(m, dtype=torch.float64, device=dev).index_add_(0, kk, t["in1"].double())
                a64 = torch.zeros(m, dtype=torch.float64, device=dev).index_add_(
                    0, kk, t["in1"].double().abs())
                bound = 4 * max(1, int(np.ceil(np.log2(max(n, 2))))) * 2.0 ** -24 * a64
                err = (t["out1"][:m].double() - v64).abs()
                ok = ok and bool((err <= bound).all())
                worst_ratio = max(worst_ratio, float((err / bound.clamp_min(1e-300)).max()))
        elif k == C.JOB_INCLUSIVE_SCAN and dt == "float32":
            x64 = t["in0"].double()
            p64 = torch.cumsum(x64, 0)
            a64 = torch.cumsum(x64.abs(), 0)
            bound = 4 * max(1, int(np.ceil(np.log2(max(n, 2))))) * 2.0 ** -24 * a64
            err = (t["out0"].double() - p64).abs()
            ok = bool((err <= bound).all())
            worst_ratio = max(worst_ratio, float((err / bound.clamp_min(1e-300)).max()))
        else:
            ok = all(torch.equal(t[name], r) for name, r in zip(("out0", "out1"), ref_out[i]))
        mismatches += 0 if ok else 1
    checked = ref_out is not None
    ref_out = None  # free the per-call copies
    _log(f"batched vs per-call: {mismatches} mismatching jobs; worst float err/bound {worst_ratio:.3g}")

    bt_ms = timed(lambda: bt.launch(use_graph=True), args.steps)
    _log(f"batched: {np.mean(bt_ms):.2f} ms/step")

    # per-stage device time: eager launches timed with HIP events on the
    # batch's stream (one event pair around each phase-0 family kernel, the
    # radix passes, and each pass kernel)
    reps = max(3, args.steps)
    acc = {k: 0.0 for k in Batch.STAGES}
    for _ in range(reps):
        W.reset_all(jobs)
        bt.launch(use_graph=False)
        for k, v in bt.stage_timing().items():
            acc[k] += v / reps
    torch.cuda.synchronize()
    bt.close()
    stage_bytes = {k: 0 for k in Batch.STAGES}
    for i, j in enumerate(jobs):
        k, dt, n = j.spec.kind, j.spec.dtype, j.spec.n
        if k == C.JOB_INCLUSIVE_SCAN:
            stage_bytes["scan_" + ("i32" if dt == "int32" else "f32")] += j.algorithmic_bytes()
        elif k == C.JOB_REDUCE_BY_KEY:
            stage_bytes["reduce_by_key_" + ("i32" if dt == "int32" else "f32")] += \
                j.algorithmic_bytes(runs.get(i))
        elif k == C.JOB_MERGE_BY_KEY:
            stage_bytes["merge_by_key"] += j.algorithmic_bytes()
        elif Batch.sort_passes(n) == 0:
            stage_bytes["sort_block"] += j.algorithmic_bytes()
        else:
            # compulsory: every key and value read and written once (2*(4+4)*n);
            # each of the 4 downsweep launches moves exactly that, each
            # upsweep reads the keys (4*n)
            stage_bytes["sort_radix"] += j.algorithmic_bytes()
            stage_bytes["sort_downsweep_kernels"] += 4 * 16 * n
            stage_bytes["sort_upsweep_kernels"] += 4 * 4 * n
    radix_launches = 4
    stages = {k: {"ms": acc[k], "bytes": stage_bytes[k],
                  "gbs": stage_bytes[k] / (acc[k] * 1e-3) / 1e9 if acc[k] > 0 else None}
              for k in Batch.STAGES}
    for k, v in stages.items():
        if v["bytes"]:
            _log(f"stage {k:18s} {v['ms']:7.3f} ms {v['gbs'] or 0:7.0f} GB/s")
    # roofline: the kernel with the largest device time per step -- a
    # phase-0 family kernel (one launch) or the radix downsweep or upsweep
    # kernel (one launch per radix pass); per-launch bytes and duration are
    # the step's averages
    kernel_of = {k: f"batch_phase0_kernel<{i}>" for i, k in enumerate(Batch.STAGES[:7])}
    kernel_of["sort_downsweep_kernels"] = "batch_radix_downsweep_kernel"
    kernel_of["sort_upsweep_kernels"] = "batch_radix_upsweep_kernel"
    dom = max(kernel_of, key=lambda k: acc[k])
    dom_launches = radix_launches if dom in ("sort_downsweep_kernels", "sort_upsweep_kernels") else 1
    dom_gbs = stages[dom]["gbs"]

    # ---------------------------------------------------------------- serving
    serve = None
    if not args.no_serve:
        from bulk_amd.serve import Server, SERVE_PER_CALL
        budget = int(args.budget_gb * 1e9)

        def serve_run(flags, arrivals_us):
            W.reset_all(jobs)
            torch.cuda.synchronize()
            s = Server(budget, max_inflight=3, max_batch_jobs=512, max_batch_elements=1 << 27,
                       flags=flags)
            # arrivals start after every request has been handed over, so the
            # client's submission speed does not shape the arrival process
            t_start = s.now_us() + 300000.0
            ids = []
            for j, off in zip(jobs, arrivals_us):
                ids.append(s.submit(j.job, j.footprint_bytes(), arrival_us=t_start + off))
            s.drain()
            times = [s.times(r) for r in ids]
            st = s.stats()
            s.close()
            lat = [(d - a) * 1e-3 for a, _, d in times]  # ms
            span_ms = (max(d for _, _, d in times) - t_start) * 1e-3
            return {"p50_ms": _pct(lat, 50), "p99_ms": _pct(lat, 99), "span_ms": span_ms,
                    "batches": st["batches"], "peak_admitted_gb": st["peak_bytes"] / 1e9}

        zero = np.zeros(len(jobs))
        serve_run(0, zero)  # warm the slots' graphs
        sat = serve_run(0, zero)
        sat_melem = total / (sat["span_ms"] * 1e-3) / 1e6
        rate = args.load * len(jobs) / (sat["span_ms"] * 1e-3)
        arr = W.poisson_arrivals(len(jobs), rate, seed=args.seed + 1)
        pois = serve_run(0, arr)
        pois_pc = serve_run(SERVE_PER_CALL, arr)
        # latency against offered load: Poisson arrivals at fractions of the
        # saturated rate
        sweep = []
        for load in (0.25, 0.5, 0.75, 0.9, 1.0):
            r = load * len(jobs) / (sat["span_ms"] * 1e-3)
            res = serve_run(0, W.poisson_arrivals(len(jobs), r, seed=args.seed + 2))
            sweep.append({"load": load, "rate_jobs_per_s": r, "p50_ms": res["p50_ms"],
                          "p99_ms": res["p99_ms"], "batches": res["batches"]})
            _log(f"poisson load {load:.2f}: p50 {res['p50_ms']:.3f} ms p99 {res['p99_ms']:.3f} ms")
        sat_pc = serve_run(SERVE_PER_CALL, zero)
        # Host-staged requests (data starts and ends in pinned host memory):
        # a sample of the stream served at once, PCIe copies included.
        from bulk_amd.batch import Job as HostJob
        W.reset_all(jobs)
        torch.cuda.synchronize()
        hsample = jobs[:min(len(jobs), 512)]
        hjobs = [HostJob(jd.job.kind, jd.job.dtype, jd.job.value_dtype, jd.job.n0, jd.job.n1,
                         {k: v.cpu().pin_memory() for k, v in jd.job.tensors.items()})
                 for jd in hsample]

        def host_run():
            for jd, h in zip(hsample, hjobs):  # sort jobs sort the host buffers in place
                if jd.restore is not None:
                    h.tensors["out0"].copy_(jd.restore[2])
                    h.tensors["out1"].copy_(jd.restore[3])
            s = Server(budget, max_inflight=3, max_batch_jobs=512, max_batch_elements=1 << 27)
            t_start = s.now_us()
            sub = [s.submit_host(h) for h in hjobs]
            s.drain()
            times = [s.times(r) for r, _ in sub]
            s.close()
            span = (max(d for _, _, d in times) - t_start) * 1e-3
            moved = sum(c * (2 if h.kind == C.JOB_SORT_BY_KEY else 1)
                        for (_, c), h in zip(sub, hjobs))
            return span, moved

        host_run()  # warm
        h_span, h_moved = host_run()
        h_elems = sum(jd.spec.n for jd in hsample)
        host_staged = {"jobs": len(hsample), "elements": h_elems, "span_ms": h_span,
                       "melem_per_s": h_elems / (h_span * 1e-3) / 1e6,
                       "host_bytes_moved": h_moved,
                       "host_gbs": h_moved / (h_span * 1e-3) / 1e9}
        del hjobs
        _log(f"served host-staged ({len(hsample)} jobs) {host_staged['melem_per_s']:.0f} Melem/s, "
             f"{host_staged['host_gbs']:.1f} GB/s over PCIe")
        serve = {
            "host_staged": host_staged,
            "saturated": {**sat, "melem_per_s": sat_melem},
            "saturated_per_call": {**sat_pc, "melem_per_s": total / (sat_pc["span_ms"] * 1e-3) / 1e6},
            "poisson_rate_jobs_per_s": rate,
            "poisson": pois,
            "poisson_per_call": pois_pc,
            "poisson_sweep": sweep,
        }
        _log(f"served saturated {sat_melem:.0f} Melem/s p99 {sat['p99_ms']:.2f} ms; "
             f"poisson p50 {pois['p50_ms']:.3f} p99 {pois['p99_ms']:.3f} ms; "
             f"per-call poisson p99 {pois_pc['p99_ms']:.3f} ms")

    # ------------------------------------------------------------ CPU baseline
    cpu = None
    if not args.no_cpu:
        import oracle as O
        done_elems, cpu_s, sample = 0, 0.0, 0
        kept, kept_bytes = [], 0  # pristine host inputs for the all-cores leg
        order = np.random.default_rng(args.seed + 3).permutation(len(jobs))
        for i in order:
            if cpu_s >= args.cpu_seconds:
                break
            j = jobs[int(i)]
            t = j.job.tensors
            j.reset()
            host = {k: v.cpu().numpy() for k, v in t.items() if k.startswith("in") or
                    (j.spec.kind == C.JOB_SORT_BY_KEY and k in ("out0", "out1"))}
            if kept_bytes < 8e9:
                kept.append((j.spec.kind, {k: v.copy() for k, v in host.items()}, j.elements))
                kept_bytes += sum(v.nbytes for v in host.values())
            a = time.perf_counter()
            if j.spec.kind == C.JOB_INCLUSIVE_SCAN:
                O.inclusive_scan(host["in0"])
            elif j.spec.kind == C.JOB_REDUCE_BY_KEY:
                O.reduce_by_key(host["in0"], host["in1"])
            elif j.spec.kind == C.JOB_MERGE_BY_KEY:
                O.merge_by_key(host["in0"], host["in1"], host["in2"], host["in3"])
            else:
                O.stable_sort_by_key(host["out0"], host["out1"])
            cpu_s += time.perf_counter() - a
            done_elems += j.elements
            sample += 1
        cpu = {"value": done_elems / cpu_s / 1e6, "unit": "Melem/s", "cores": 1, "kind": "port",
               "sample": f"{sample} of the {len(jobs)} jobs ({done_elems} elements, {cpu_s:.1f} s), "
                         "random order, sequential restatement of the reference algorithm bodies"}
        _log(f"cpu baseline {cpu['value']:.1f} Melem/s over {sample} jobs")

        # The same bodies on every host core: one job per thread at a time
        # (the oracle's ctypes calls release the GIL).
        from concurrent.futures import ThreadPoolExecutor

        def one(item):
            kind, h, _ = item
            if kind == C.JOB_INCLUSIVE_SCAN:
                O.inclusive_scan(h["in0"])
            elif kind == C.JOB_REDUCE_BY_KEY:
                O.reduce_by_key(h["in0"], h["in1"])
            elif kind == C.JOB_MERGE_BY_KEY:
                O.merge_by_key(h["in0"], h["in1"], h["in2"], h["in3"])
            else:  # sorts in place: sort a copy so the inputs stay unsorted
                O.stable_sort_by_key(h["out0"].copy(), h["out1"].copy())

        # 16 threads: the host cores a single-GPU share of the machine gets
        # (one GPU of eight on a node); fewer when fewer are visible
        threads = min(16, len(os.sched_getaffinity(0)))
        with ThreadPoolExecutor(threads) as ex:
            a = time.perf_counter()
            list(ex.map(one, kept))
            wall = time.perf_counter() - a
        kept_elems = sum(e for _, _, e in kept)
        cpu["all_cores"] = {"value": kept_elems / wall / 1e6, "unit": "Melem/s", "cores": threads,
                            "sample": f"{len(kept)} of those jobs ({kept_elems} elements, "
                                      f"{wall:.2f} s wall), one job per thread"}
        del kept
        _log(f"cpu baseline on {threads} threads {cpu['all_cores']['value']:.1f} Melem/s")

        # The plain int32 inclusive scan of 2^24 elements (uniform [-1000, 1000],
        # the single-op table's input) through the sequential body on one
        # core, beside the GPU kernel's result, which must equal it bit for bit.
        from bulk_amd import api as A
        g1 = torch.Generator(device=dev)
        g1.manual_seed(0)
        x1 = torch.randint(-1000, 1001, (1 << 24,), generator=g1, device=dev, dtype=torch.int32)
        gpu1 = A.inclusive_scan(x1).cpu().numpy()
        h1 = x1.cpu().numpy()
        ts = []
        for _ in range(6):
            a = time.perf_counter()
            ref1 = O.inclusive_scan(h1)
            ts.append(time.perf_counter() - a)
        t1 = float(np.median(ts[1:]))
        cpu["scan_int32_2^24"] = {"ms": t1 * 1e3, "melem_per_s": (1 << 24) / t1 / 1e6, "cores": 1,
                                  "gpu_equals_cpu": bool(np.array_equal(gpu1, ref1))}
        # The same scan on `threads` threads through the reference's own
        # three-phase decomposition (scan.cu:109-145: per-partition sums,
        # a scan of the carries, per-partition scans seeded by them).
        ts = []
        for _ in range(6):
            a = time.perf_counter()
            mt1 = O.mt_inclusive_scan_i32(h1, 0, threads)
            ts.append(time.perf_counter() - a)
        tm = float(np.median(ts[1:]))
        cpu["scan_int32_2^24_threads"] = {"ms": tm * 1e3, "melem_per_s": (1 << 24) / tm / 1e6,
                                          "cores": threads,
                                          "equals_one_thread": bool(np.array_equal(mt1, ref1))}
        cpu["host_cpu"] = _cpu_model()
        del x1, gpu1, h1, ref1, mt1
        _log(f"cpu int32 scan 2^24: {t1 * 1e3:.2f} ms on 1 thread, {tm * 1e3:.2f} ms on {threads}; "
             f"GPU result equal: {cpu['scan_int32_2^24']['gpu_equals_cpu']}")

    # ------------------------------------------------- single ops at full size
    ops = None
    if not args.no_ops:
        from bulk_amd import opbench
        del jobs  # frees the stream's device memory before the 2^28 inputs
        torch.cuda.empty_cache()
        ops = opbench.run(iters=10, log=_log)
        jobs = None

    # ------------------------------------------------------------------ report
    bt_mean = float(np.mean(bt_ms))
    value = total / (bt_mean * 1e-3) / 1e6
    pc_mean = float(np.mean(pc_ms)) if pc_ms else None
    pc_value = total / (pc_mean * 1e-3) / 1e6 if pc_ms else None
    tr = _read_traffic(args.traffic_file, kernel_of[dom])
    traffic = None
    if tr is not None and tr[1] == njobs and tr[2] == args.seed:
        traffic = tr[0]
    line = {
        "metric": "batched vs per-call Melem/s + p99 latency, 4k mixed jobs; GB/s %HBM peak",
        "value": value,
        "unit": "Melem/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": bt_mean,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32+float32",
        "data": "synthetic",
        "config": {"workload": f"mixed-job stream: {njobs} jobs (scan 40%, reduce_by_key 30%, "
                               "merge 20%, sort_by_key 10%), log-uniform 2^10..2^22 elements, "
                               "int32/float32", "jobs": njobs, "elements": total,
                   "seed": args.seed, "hbm_budget_gb": args.budget_gb},
        "per_call_melem_per_s": pc_value,
        "per_call_ms_per_step": pc_mean,
        "speedup_vs_per_call": value / pc_value if pc_value else None,
        "batched_equals_per_call": (mismatches == 0) if checked else None,
        "float_err_over_bound_max": worst_ratio,
        "serve": serve,
        "p99_ms": serve["poisson"]["p99_ms"] if serve else None,
        "p50_ms": serve["poisson"]["p50_ms"] if serve else None,
        "copy_gbs": copy_gbs,
        "hbm_ceiling": ceiling,
        "roofline": {"bound": "hbm", "kernel": f"{kernel_of[dom]} ({dom})",
                     "achieved": dom_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dom_gbs / HBM_PEAK_GBS, "frac_of_copy": dom_gbs / copy_gbs,
                     "traffic": traffic, "algorithmic_bytes": stages[dom]["bytes"] / dom_launches,
                     "avg_launch_ms": stages[dom]["ms"] / dom_launches,
                     "launches_per_step": dom_launches},
        "stages": stages,
        "ops": ops,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
