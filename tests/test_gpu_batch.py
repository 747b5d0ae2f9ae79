"""GPU parity of the batch executor: a mixed job list run as one persistent
launch (eagerly and as a replayed hipGraph) and as one plain call per job,
each job's result checked against the oracle."""
import numpy as np
import pytest
import torch

from bulk_amd import _capi as C

pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _make_jobs(orc, sizes, seed=0):
    """Jobs over every supported kind; returns (jobs, checks) where each check
    compares a job's outputs with the oracle."""
    from bulk_amd import batch as B
    jobs, checks = [], []
    rng = np.random.default_rng(seed)
    for i, n in enumerate(sizes):
        kind = i % 6
        if kind == 0:  # int32 scan
            x = orc.gen_uniform_i32(seed + i, -1000, 1000, n)
            dx, out = _dev(x), torch.empty(n, dtype=torch.int32, device="cuda")
            jobs.append(B.scan_job(dx, out))
            ref = orc.inclusive_scan(x)
            checks.append(lambda out=out, ref=ref: np.array_equal(out.cpu().numpy(), ref))
        elif kind == 1:  # float32 scan, within the stated bound
            x = orc.gen_uniform_f32(seed + i, n)
            dx, out = _dev(x), torch.empty(n, dtype=torch.float32, device="cuda")
            jobs.append(B.scan_job(dx, out))
            hi, lo, ab = orc.prefix_sum_exact(x)
            bound = orc.float_tolerance(n, 2.0 ** -24, ab)
            checks.append(lambda out=out, hi=hi, lo=lo, b=bound:
                          bool(np.all(np.abs(out.cpu().numpy().astype(np.float64) - hi - lo) <= b)))
        elif kind == 2:  # reduce_by_key, int32 keys, int64 or float32 values
            keys, runs = orc.gen_runs_geometric(seed + i, n, 8.0)
            if i % 12 == 2:
                vals = orc.gen_uniform_i64(seed + i, -(1 << 20), 1 << 20, n)
            elif i % 18 == 8:
                vals = orc.gen_uniform_f32(seed + i, n, -1.0, 1.0)
            else:
                vals = orc.gen_uniform_i32(seed + i, -50, 50, n)
            dk, dv = _dev(keys), _dev(vals)
            ko = torch.empty(n, dtype=torch.int32, device="cuda")
            vo = torch.empty(n, dtype=dv.dtype, device="cuda")
            cnt = torch.full((1,), -1, dtype=torch.int64, device="cuda")
            jobs.append(B.reduce_by_key_job(dk, dv, ko, vo, cnt))
            if vals.dtype == np.float32:  # keys and count exact, sums within the bound
                v64 = vals.astype(np.float64)
                starts = np.flatnonzero(np.r_[True, keys[1:] != keys[:-1]])
                exact = np.add.reduceat(v64, starts)
                lens = np.diff(np.r_[starts, n])
                bound = 4 * np.maximum(1, np.ceil(np.log2(np.maximum(lens, 2)))) * 2.0 ** -24 * \
                    np.add.reduceat(np.abs(v64), starts)
                rk = keys[starts + lens - 1]
                checks.append(lambda ko=ko, vo=vo, cnt=cnt, rk=rk, ex=exact, bd=bound:
                              int(cnt.item()) == rk.size
                              and np.array_equal(ko.cpu().numpy()[:rk.size], rk)
                              and bool(np.all(np.abs(vo.cpu().numpy()[:rk.size].astype(np.float64) - ex)
                                              <= bd)))
            else:
                rk, rv = orc.reduce_by_key(keys, vals, textbook=True)
                checks.append(lambda ko=ko, vo=vo, cnt=cnt, rk=rk, rv=rv:
                              int(cnt.item()) == rk.size
                              and np.array_equal(ko.cpu().numpy()[:rk.size], rk)
                              and np.array_equal(vo.cpu().numpy()[:rk.size], rv))
        elif kind == 3:  # merge_by_key, float32 keys with ties, int32 payload
            pool = orc.gen_uniform_f32(seed + 99, 64)
            n1 = n // 2 + 1
            a = orc.gen_merge_keys(seed + i, 0, n1, pool, 0.1)
            b = orc.gen_merge_keys(seed + i, 1, n - n1 + 3, pool, 0.1)
            va = np.arange(a.size, dtype=np.int32)
            vb = (1 << 28) + np.arange(b.size, dtype=np.int32)
            ko = torch.empty(a.size + b.size, dtype=torch.float32, device="cuda")
            vo = torch.empty(a.size + b.size, dtype=torch.int32, device="cuda")
            jobs.append(B.merge_by_key_job(_dev(a), _dev(va), _dev(b), _dev(vb), ko, vo))
            rk, rv = orc.merge_by_key(a, va, b, vb)
            checks.append(lambda ko=ko, vo=vo, rk=rk, rv=rv:
                          np.array_equal(ko.cpu().numpy(), rk) and np.array_equal(vo.cpu().numpy(), rv))
        elif kind == 4:  # stable sort by key, int32 keys with ties, index payload
            keys = orc.gen_uniform_i32(seed + i, 0, max(2, n // 16), n)
            vals = np.arange(n, dtype=np.int32)
            dk, dv = _dev(keys), _dev(vals)
            jobs.append(B.sort_by_key_job(dk, dv))
            order = np.argsort(keys, kind="stable")
            rk, rv = keys[order], vals[order]
            keys0, vals0 = dk.clone(), dv.clone()
            checks.append(lambda dk=dk, dv=dv, rk=rk, rv=rv:
                          np.array_equal(dk.cpu().numpy(), rk) and np.array_equal(dv.cpu().numpy(), rv))
            # sort is in place: keep the input to restore between runs
            jobs[-1].restore = (dk, dv, keys0, vals0)
        else:  # reduce, int32
            x = orc.gen_uniform_i32(seed + i, -1000, 1000, n)
            out = torch.zeros(1, dtype=torch.int32, device="cuda")
            jobs.append(B.reduce_job(_dev(x), out))
            ref = int(orc.reduce(x, 0))
            checks.append(lambda out=out, ref=ref: int(out.item()) == ref)
    return jobs, checks


def _restore(jobs):
    for j in jobs:
        r = getattr(j, "restore", None)
        if r is not None:
            r[0].copy_(r[2])
            r[1].copy_(r[3])


SIZES = [1, 1000, 4095, 4096, 4097, 12345, 16384, 16385, 65537, 200001, 3, 1 << 18,
         77777, 8191, 8193, 1 << 20, 50000, 30000, 2, 4096 * 3 + 1, 123457, 9999, 1 << 16, 7]


@pytest.mark.parametrize("mode", ["graph", "eager", "branches", "per_call"])
def test_batch_matches_oracle(bx, orc, mode):
    from bulk_amd.batch import Batch
    jobs, checks = _make_jobs(orc, SIZES, seed=1)
    b = Batch(max_jobs=64, max_elements=1 << 23)
    b.set_jobs(jobs)
    print(f"batch mode={mode} jobs={len(jobs)}", flush=True)
    if mode == "per_call":
        b.launch_per_call()
    else:
        b.launch({"graph": 1, "eager": 0, "branches": 2}[mode])
    torch.cuda.synchronize()
    bad = [i for i, c in enumerate(checks) if not c()]
    assert bad == [], f"jobs {bad} differ (sizes {[SIZES[i] for i in bad]})"
    b.close()


@pytest.mark.parametrize("shape", [0, 1])
@pytest.mark.parametrize("mode", ["graph", "eager"])
def test_batch_family_shapes(bx, orc, shape, mode):
    """Every family in each of its tile shapes (forced through the tuning
    table before the batch is created) matches the oracle."""
    from bulk_amd.batch import Batch
    keys = [(C.ALGO_BATCH_SCAN, C.INT32), (C.ALGO_BATCH_SCAN, C.FLOAT32),
            (C.ALGO_BATCH_REDUCE_BY_KEY, C.INT32), (C.ALGO_BATCH_REDUCE_BY_KEY, C.FLOAT32),
            (C.ALGO_BATCH_REDUCE_BY_KEY, C.INT64), (C.ALGO_BATCH_MERGE, C.INT32)]
    try:
        for algo, dt in keys:
            C.check(C.lib().bulkx_set_tuning(algo, dt, 0, shape), "bulkx_set_tuning")
        jobs, checks = _make_jobs(orc, SIZES + [3 * 16384 + 5, 5 * 8192 + 1], seed=7)
        b = Batch(max_jobs=64, max_elements=1 << 23)
        got = b.family_shapes()
        assert all(got[f][0] == shape for f in ("scan_i32", "scan_f32", "reduce_by_key_i32",
                                                 "reduce_by_key_f32", "reduce_by_key_i64",
                                                 "merge_by_key")), got
        b.set_jobs(jobs)
        print(f"batch shape={shape} mode={mode} tiles={got}", flush=True)
        b.launch(use_graph=(mode == "graph"))
        torch.cuda.synchronize()
        bad = [i for i, c in enumerate(checks) if not c()]
        assert bad == [], f"jobs {bad} differ"
        assert b.status() == 0
        b.close()
    finally:
        C.check(C.lib().bulkx_load_tuning(None, None), "bulkx_load_tuning")


def test_batch_graph_replays_and_new_job_lists(bx, orc):
    """Replays of one captured graph, with eager work between them and a new
    job list uploaded before the last replays, all match the oracle."""
    from bulk_amd.batch import Batch
    b = Batch(max_jobs=64, max_elements=1 << 23)
    # every round count's graph recorded up front; launches then only replay
    C.check(C.lib().bulkx_batch_prepare_graphs(b._h), "bulkx_batch_prepare_graphs")
    jobs, checks = _make_jobs(orc, SIZES, seed=2)
    b.set_jobs(jobs)
    for rep in range(3):
        _restore(jobs)
        for j in jobs:  # scribble over the outputs eagerly
            for name in ("out0", "out1"):
                t = j.tensors.get(name)
                if t is not None and j.kind != C.JOB_SORT_BY_KEY:
                    t.fill_(7)
        b.launch(use_graph=True)
        torch.cuda.synchronize()
        assert all(c() for c in checks), f"replay {rep}"
    jobs2, checks2 = _make_jobs(orc, list(reversed(SIZES))[:17], seed=3)
    b.set_jobs(jobs2)
    for rep in range(2):
        _restore(jobs2)
        b.launch(use_graph=True)
        torch.cuda.synchronize()
        assert all(c() for c in checks2), f"new list replay {rep}"
    b.close()


def test_batch_rejects_over_capacity(bx, orc):
    from bulk_amd import batch as B
    from bulk_amd._capi import BulkError
    b = B.Batch(max_jobs=2, max_elements=1000)
    x = torch.zeros(2000, dtype=torch.int32, device="cuda")
    with pytest.raises(BulkError):
        b.set_jobs([B.scan_job(x, x)])
    with pytest.raises(BulkError):
        b.set_jobs([B.scan_job(x[:10], x[:10])] * 3)
    b.close()


@pytest.mark.parametrize("mode", ["graph", "eager"])
def test_batch_reduce_by_key_long_runs(bx, orc, mode):
    """Runs longer than a tile (tiles with no head carry the open run through
    the aggregates), a single run over a whole job, and float values within
    the bound, in one batch."""
    from bulk_amd import batch as B
    jobs, checks = [], []
    for i, (n, run) in enumerate([(100003, 20000), (70000, 70000), (300001, 9000), (5000, 1)]):
        keys = (np.arange(n) // run).astype(np.int32)
        vals = orc.gen_uniform_i64(40 + i, -(1 << 20), 1 << 20, n)
        dk, dv = _dev(keys), _dev(vals)
        ko = torch.empty(n, dtype=torch.int32, device="cuda")
        vo = torch.empty(n, dtype=torch.int64, device="cuda")
        cnt = torch.full((1,), -1, dtype=torch.int64, device="cuda")
        jobs.append(B.reduce_by_key_job(dk, dv, ko, vo, cnt))
        rk, rv = orc.reduce_by_key(keys, vals, textbook=True)
        checks.append(lambda ko=ko, vo=vo, cnt=cnt, rk=rk, rv=rv:
                      int(cnt.item()) == rk.size and np.array_equal(ko.cpu().numpy()[:rk.size], rk)
                      and np.array_equal(vo.cpu().numpy()[:rk.size], rv))
        f = orc.gen_uniform_f32(50 + i, n, -1.0, 1.0)
        fo = torch.empty(n, dtype=torch.float32, device="cuda")
        cnt2 = torch.full((1,), -1, dtype=torch.int64, device="cuda")
        ko2 = torch.empty(n, dtype=torch.int32, device="cuda")
        jobs.append(B.reduce_by_key_job(dk, _dev(f), ko2, fo, cnt2))
        f64 = f.astype(np.float64)
        starts = np.flatnonzero(np.r_[True, keys[1:] != keys[:-1]])
        exact = np.add.reduceat(f64, starts)
        lens = np.diff(np.r_[starts, n])
        # 4 * ceil(log2 len) * 2^-24 * sum|x| per run
        bound = 4 * np.maximum(1, np.ceil(np.log2(np.maximum(lens, 2)))) * 2.0 ** -24 * \
            np.add.reduceat(np.abs(f64), starts)
        checks.append(lambda fo=fo, cnt2=cnt2, exact=exact, bound=bound:
                      int(cnt2.item()) == exact.size and
                      bool(np.all(np.abs(fo.cpu().numpy()[:exact.size].astype(np.float64) - exact) <= bound)))
    b = B.Batch(max_jobs=16, max_elements=1 << 21)
    b.set_jobs(jobs)
    print(f"batch long runs mode={mode}", flush=True)
    for rep in range(2):
        b.launch(use_graph=(mode == "graph"))
        torch.cuda.synchronize()
        bad = [k for k, c in enumerate(checks) if not c()]
        assert bad == [], f"replay {rep}: jobs {bad} differ"
    b.close()


@pytest.mark.parametrize("mode", ["graph", "eager"])
def test_batch_reduce_by_key_tile_edges(bx, orc, mode):
    """Run layouts at the edges of the compact form's per-tile slots: every
    element its own run over many tiles (each tile fills all its slots),
    runs exactly one tile long (every tile starts with a head and ends one
    run), half-tile runs with a one-element last tile, runs one short of a
    tile (the boundary moves through the tiles), a job of exactly one tile
    and one of a tile plus one element; int32, int64 and float32 values."""
    from bulk_amd import batch as B
    probe = B.Batch(max_jobs=1, max_elements=1)
    T = probe.family_shapes()["reduce_by_key_i32"][1]
    probe.close()
    layouts = [(3 * T + 5, 1), (4 * T, T), (4 * T + 1, T // 2), (5 * T + 17, T - 1), (T, 7), (T + 1, 2)]
    jobs, checks = [], []
    for i, (n, run) in enumerate(layouts):
        keys = (np.arange(n) // run).astype(np.int32)
        for vk in ("i32", "i64", "f32"):
            if vk == "i32":
                vals = orc.gen_uniform_i32(70 + i, -50, 50, n)
            elif vk == "i64":
                vals = orc.gen_uniform_i64(80 + i, -(1 << 40), 1 << 40, n)
            else:
                vals = orc.gen_uniform_f32(90 + i, n, -1.0, 1.0)
            dk, dv = _dev(keys), _dev(vals)
            ko = torch.empty(n, dtype=torch.int32, device="cuda")
            vo = torch.empty(n, dtype=dv.dtype, device="cuda")
            cnt = torch.full((1,), -1, dtype=torch.int64, device="cuda")
            jobs.append(B.reduce_by_key_job(dk, dv, ko, vo, cnt))
            if vk == "f32":
                v64 = vals.astype(np.float64)
                starts = np.flatnonzero(np.r_[True, keys[1:] != keys[:-1]])
                exact = np.add.reduceat(v64, starts)
                lens = np.diff(np.r_[starts, n])
                bound = 4 * np.maximum(1, np.ceil(np.log2(np.maximum(lens, 2)))) * 2.0 ** -24 * \
                    np.add.reduceat(np.abs(v64), starts)
                rk = keys[starts + lens - 1]
                checks.append((f"{n}/{run}/{vk}", lambda ko=ko, vo=vo, cnt=cnt, rk=rk, ex=exact, bd=bound:
                               int(cnt.item()) == rk.size
                               and np.array_equal(ko.cpu().numpy()[:rk.size], rk)
                               and bool(np.all(np.abs(vo.cpu().numpy()[:rk.size].astype(np.float64) - ex)
                                               <= bd))))
            else:
                rk, rv = orc.reduce_by_key(keys, vals, textbook=True)
                checks.append((f"{n}/{run}/{vk}", lambda ko=ko, vo=vo, cnt=cnt, rk=rk, rv=rv:
                               int(cnt.item()) == rk.size
                               and np.array_equal(ko.cpu().numpy()[:rk.size], rk)
                               and np.array_equal(vo.cpu().numpy()[:rk.size], rv)))
    b = B.Batch(max_jobs=len(jobs), max_elements=sum(j.elements for j in jobs))
    b.set_jobs(jobs)
    print(f"batch reduce_by_key tile edges mode={mode} tile={T} jobs={len(jobs)}", flush=True)
    b.launch(1 if mode == "graph" else 0)
    torch.cuda.synchronize()
    assert b.status() == 0
    bad = [name for name, ok in checks if not ok()]
    b.close()
    assert bad == [], bad
