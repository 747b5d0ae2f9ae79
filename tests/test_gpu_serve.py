"""GPU tests of the serving loop and the mixed-job workload: each result is
the same alone and in the mix (integer work bit-exact, float work within
4*ceil(log2 n)*2^-24*sum|x| of FP64), a sample matches the oracle, admission
respects the HBM budget, and batches of different slots overlap."""
import numpy as np
import pytest
import torch

from bulk_amd import _capi as C

pytestmark = pytest.mark.gpu


def _small_stream(num_jobs=192, seed=11):
    from bulk_amd import workload as W
    specs = W.make_specs(num_jobs, seed=seed, log2_min=10, log2_max=17)
    return W.build_jobs(specs, seed=seed + 1)


def _snapshot(jd):
    t = jd.job.tensors
    if jd.spec.kind == C.JOB_REDUCE_BY_KEY:
        m = int(t["out2"].item())
        return (m, t["out0"][:m].clone(), t["out1"][:m].clone())
    if jd.spec.kind == C.JOB_SORT_BY_KEY:
        return (t["out0"].clone(), t["out1"].clone())
    return tuple(t[k].clone() for k in ("out0", "out1") if k in t)


def _bound(n, absum):
    return 4 * max(1, int(np.ceil(np.log2(max(n, 2))))) * 2.0 ** -24 * absum


def _float_ok(jd, out):
    """Float results against FP64, within the stated bound."""
    t = jd.job.tensors
    n = jd.spec.n
    if jd.spec.kind == C.JOB_INCLUSIVE_SCAN:
        x = t["in0"].double()
        err = (out[0].double() - torch.cumsum(x, 0)).abs()
        return bool((err <= _bound(n, torch.cumsum(x.abs(), 0))).all())
    m, _, v = out
    kk = t["in0"].long()
    v64 = torch.zeros(m, dtype=torch.float64, device="cuda").index_add_(0, kk, t["in1"].double())
    a64 = torch.zeros(m, dtype=torch.float64, device="cuda").index_add_(0, kk, t["in1"].double().abs())
    return bool(((v.double() - v64).abs() <= _bound(n, a64)).all())


def _same(jd, a, b):
    is_float = jd.spec.dtype == "float32" and jd.spec.kind in (C.JOB_INCLUSIVE_SCAN,
                                                               C.JOB_REDUCE_BY_KEY)
    if jd.spec.kind == C.JOB_REDUCE_BY_KEY:
        if a[0] != b[0] or not torch.equal(a[1], b[1]):
            return False
        return _float_ok(jd, a) and _float_ok(jd, b) if is_float else torch.equal(a[2], b[2])
    if is_float:
        return _float_ok(jd, a) and _float_ok(jd, b)
    return all(torch.equal(x, y) for x, y in zip(a, b))


def _run_alone(jobs):
    """Every job through the plain entry points, one at a time."""
    from bulk_amd.batch import Batch
    out = []
    b = Batch(max_jobs=1, max_elements=1 << 24)
    for jd in jobs:
        jd.reset()
        b.set_jobs([jd.job])
        b.launch_per_call()
        torch.cuda.synchronize()
        out.append(_snapshot(jd))
    b.close()
    return out


def test_served_mix_equals_alone(bx, orc):
    from bulk_amd import workload as W
    from bulk_amd.serve import Server
    jobs = _small_stream()
    alone = _run_alone(jobs)
    W.reset_all(jobs)
    for jd in jobs:  # scribble over outputs so stale results cannot pass
        for name in ("out0", "out1"):
            if jd.spec.kind != C.JOB_SORT_BY_KEY and name in jd.job.tensors:
                jd.job.tensors[name].fill_(3)
    torch.cuda.synchronize()
    budget = 24 << 20  # forces admission to wait for memory
    biggest = max(jd.footprint_bytes() for jd in jobs)
    assert biggest <= budget
    srv = Server(budget, max_inflight=3, max_batch_jobs=32, max_batch_elements=1 << 20)
    arr = W.poisson_arrivals(len(jobs), rate_per_s=40000.0, seed=3)
    t0 = srv.now_us() + 1000.0
    ids = [srv.submit(jd.job, jd.footprint_bytes(), arrival_us=t0 + a) for jd, a in zip(jobs, arr)]
    srv.drain()
    torch.cuda.synchronize()
    st = srv.stats()
    times = [srv.times(r) for r in ids]
    srv.close()
    print(f"served {st}", flush=True)
    assert st["completed"] == len(jobs)
    assert st["batches"] > 1
    assert st["peak_bytes"] <= budget
    for a, l, d in times:
        assert a <= l + 1.0 and l <= d
    bad = [i for i, jd in enumerate(jobs) if not _same(jd, alone[i], _snapshot(jd))]
    assert bad == [], f"jobs {bad} differ alone vs in the mix"


def test_served_batches_overlap_and_match_oracle(bx, orc):
    """All requests at once: slots run concurrently (launch of one batch
    before the completion of another), and a sample of jobs equals the oracle."""
    from bulk_amd import workload as W
    from bulk_amd.serve import Server
    jobs = _small_stream(num_jobs=160, seed=21)
    W.reset_all(jobs)
    torch.cuda.synchronize()  # the server's streams do not wait on torch's stream
    srv = Server(1 << 34, max_inflight=3, max_batch_jobs=16, max_batch_elements=1 << 19)
    ids = [srv.submit(jd.job, jd.footprint_bytes()) for jd in jobs]
    srv.drain()
    torch.cuda.synchronize()
    times = np.array([srv.times(r) for r in ids])
    srv.close()
    launches = sorted(set(times[:, 1].tolist()))
    overlap = any(((times[:, 1] < l) & (times[:, 2] > l)).any() for l in launches)
    assert overlap, "no batch was launched while another was in flight"
    rng = np.random.default_rng(0)
    for i in rng.choice(len(jobs), size=24, replace=False):
        jd = jobs[int(i)]
        t = {k: v.cpu().numpy() for k, v in jd.job.tensors.items()}
        k = jd.spec.kind
        if k == C.JOB_INCLUSIVE_SCAN:
            if jd.spec.dtype == "int32":
                assert np.array_equal(t["out0"], orc.inclusive_scan(t["in0"]))
            else:
                hi, lo, ab = orc.prefix_sum_exact(t["in0"])
                b = orc.float_tolerance(jd.spec.n, 2.0 ** -24, ab)
                assert np.all(np.abs(t["out0"].astype(np.float64) - hi - lo) <= b)
        elif k == C.JOB_REDUCE_BY_KEY:
            rk, rv = orc.reduce_by_key(t["in0"], t["in1"], textbook=True)
            m = int(t["out2"][0])
            assert m == rk.size and np.array_equal(t["out0"][:m], rk)
            if jd.spec.dtype == "int32":
                assert np.array_equal(t["out1"][:m], rv)
        elif k == C.JOB_MERGE_BY_KEY:
            rk, rv = orc.merge_by_key(t["in0"], t["in1"], t["in2"], t["in3"])
            assert np.array_equal(t["out0"], rk) and np.array_equal(t["out1"], rv)
        else:
            k0, v0 = jd.restore[2].cpu().numpy(), jd.restore[3].cpu().numpy()
            rk, rv = orc.stable_sort_by_key(k0, v0)
            assert np.array_equal(t["out0"], rk) and np.array_equal(t["out1"], rv)


def test_server_per_call_mode_and_budget_errors(bx):
    from bulk_amd import workload as W
    from bulk_amd.serve import Server, SERVE_PER_CALL
    jobs = _small_stream(num_jobs=24, seed=31)
    alone = _run_alone(jobs)
    W.reset_all(jobs)
    torch.cuda.synchronize()
    srv = Server(1 << 30, max_inflight=2, max_batch_jobs=8, max_batch_elements=1 << 20,
                 flags=SERVE_PER_CALL)
    for jd in jobs:
        srv.submit(jd.job, jd.footprint_bytes())
    srv.drain()
    torch.cuda.synchronize()
    assert srv.stats()["batches"] == len(jobs)
    bad = [i for i, jd in enumerate(jobs) if not _same(jd, alone[i], _snapshot(jd))]
    assert bad == []
    with pytest.raises(C.BulkError):  # larger than the whole budget
        srv.submit(jobs[0].job, 1 << 31)
    srv.close()



def test_host_staged_requests_equal_alone(bx):
    """Requests over pinned host buffers: staged in and out by the server,
    under a budget that makes admission wait, each result equal to the same
    job run alone on the device (floats within the bound)."""
    from bulk_amd import workload as W
    from bulk_amd.batch import Job
    from bulk_amd.serve import Server
    jobs = _small_stream(num_jobs=64, seed=41)
    alone = _run_alone(jobs)
    W.reset_all(jobs)
    torch.cuda.synchronize()
    host = []
    for jd in jobs:
        ht = {k: v.cpu().pin_memory() for k, v in jd.job.tensors.items()}
        for name in ("out0", "out1"):  # scribble over outputs so stale results cannot pass
            if jd.spec.kind != C.JOB_SORT_BY_KEY and name in ht:
                ht[name].fill_(3)
        j = jd.job
        host.append(Job(j.kind, j.dtype, j.value_dtype, j.n0, j.n1, ht))
    budget = 8 << 20
    srv = Server(budget, max_inflight=3, max_batch_jobs=16, max_batch_elements=1 << 20)
    ids, charged = zip(*[srv.submit_host(h) for h in host])
    srv.drain()
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(jobs) and st["peak_bytes"] <= budget and st["batches"] > 1
    assert all(c > 0 and c % 256 == 0 for c in charged)  # buffers of 256-byte multiples
    scan = next(i for i, jd in enumerate(jobs) if jd.spec.kind == C.JOB_INCLUSIVE_SCAN)
    assert charged[scan] == 2 * ((jobs[scan].spec.n * 4 + 255) // 256 * 256)

    def snap(jd, ht):
        if jd.spec.kind == C.JOB_REDUCE_BY_KEY:
            m = int(ht["out2"].item())
            return (m, ht["out0"][:m].cuda(), ht["out1"][:m].cuda())
        if jd.spec.kind == C.JOB_SORT_BY_KEY:
            return (ht["out0"].cuda(), ht["out1"].cuda())
        return tuple(ht[k].cuda() for k in ("out0", "out1") if k in ht)

    bad = [i for i, jd in enumerate(jobs) if not _same(jd, alone[i], snap(jd, host[i].tensors))]
    assert bad == [], f"host-staged jobs {bad} differ from the device runs"
