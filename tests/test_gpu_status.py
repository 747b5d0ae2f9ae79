"""GPU tests of the deferred error paths: a look-back wait that gives up is
reported as BULKX_ERR_LOOKBACK (never as success) by the plain entry points,
the batch and the server, and a server that meets a bad request refuses it
or fails it without hanging.

The give-up is forced from the host: a spin limit of 0 makes every look-back
give up at once without waiting, so every kernel exits normally."""
import ctypes

import pytest
import torch

from bulk_amd import _capi as C

pytestmark = pytest.mark.gpu

DEFAULT_SPIN = 1 << 22


@pytest.fixture
def give_up(bx):
    bx.set_lookback_spin_limit(0)
    try:
        yield bx
    finally:
        bx.set_lookback_spin_limit(DEFAULT_SPIN)
        torch.cuda.synchronize()
        bx.stream_status()  # leave the sticky word clear for later tests


def _ints(n, seed=0):
    g = torch.Generator(device="cpu").manual_seed(seed)
    return torch.randint(-1000, 1001, (n,), generator=g, dtype=torch.int32).cuda()


def test_status_clear_after_normal_calls(bx):
    x = _ints(1 << 22)
    y = bx.inclusive_scan(x)
    assert bx.stream_status() == 0
    assert torch.equal(y.cpu(), torch.cumsum(x.cpu(), 0, dtype=torch.int32))


def test_scan_give_up_is_reported(give_up):
    x = _ints(1 << 22)
    give_up.inclusive_scan(x)  # multi-tile: every tile after the first looks back
    assert give_up.stream_status() == C.ERR_LOOKBACK
    # the word is cleared by the query
    assert give_up.stream_status() == 0


def test_scan_give_up_raises_through_check(give_up):
    x = _ints(1 << 21)
    give_up.inclusive_scan(x)
    with pytest.raises(C.BulkError, match="LOOKBACK"):
        give_up.check_stream()


def test_single_pass_reduce_by_key_give_up(give_up):
    """The single-pass (look-back) reduce_by_key, forced by a tuning
    override, reports the give-up; the three-pass and compact forms have no
    look-back."""
    from bulk_amd import api
    n = 1 << 18  # size class 1
    lib = C.lib()
    C.check(lib.bulkx_set_tuning(C.ALGO_REDUCE_BY_KEY, C.INT32, 1, 0), "set_tuning")
    try:
        assert api.choose_sizes(C.ALGO_REDUCE_BY_KEY, C.INT32, n)[0] > 1
        keys = torch.arange(n, dtype=torch.int32, device="cuda") // 5
        vals = torch.ones(n, dtype=torch.int32, device="cuda")
        ko, vo = torch.empty_like(keys), torch.empty_like(vals)
        cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
        api.reduce_by_key_async(keys, vals, ko, vo, cnt)
        assert give_up.stream_status() == C.ERR_LOOKBACK
        with pytest.raises(C.BulkError, match="LOOKBACK"):
            give_up.my_reduce_by_key(keys, vals)
    finally:
        ent = ctypes.c_int()
        C.check(lib.bulkx_load_tuning(None, ctypes.byref(ent)), "load_tuning")


def test_batch_give_up_is_reported(give_up):
    from bulk_amd import batch as B
    x = _ints(1 << 20)
    y = torch.empty_like(x)
    bt = B.Batch(max_jobs=4, max_elements=1 << 21)
    try:
        bt.set_jobs([B.scan_job(x, y)])
        bt.launch(use_graph=True)
        assert bt.status() == C.ERR_LOOKBACK
        bt.launch(use_graph=False)
        assert bt.status() == C.ERR_LOOKBACK
    finally:
        bt.close()


def test_batch_status_ok(bx):
    from bulk_amd import batch as B
    x = _ints(1 << 20, seed=3)
    y = torch.empty_like(x)
    bt = B.Batch(max_jobs=4, max_elements=1 << 21)
    try:
        bt.set_jobs([B.scan_job(x, y)])
        bt.launch(use_graph=True)
        assert bt.status() == 0
        assert torch.equal(y.cpu(), torch.cumsum(x.cpu(), 0, dtype=torch.int32))
    finally:
        bt.close()


def test_server_reports_give_up_and_closes(give_up):
    from bulk_amd import batch as B
    from bulk_amd import serve as S
    x = _ints(1 << 20, seed=4)
    y = torch.empty_like(x)
    srv = S.Server(hbm_budget_bytes=1 << 30, max_inflight=2, max_batch_jobs=8,
                   max_batch_elements=1 << 22)
    try:
        srv.submit(B.scan_job(x, y), 8 << 20)
        with pytest.raises(C.BulkError, match="LOOKBACK"):
            srv.drain()
    finally:
        srv.close()  # must return: the worker leaves its loop after the error


def test_server_refuses_unsupported_job_and_keeps_serving(bx):
    """A job the batch cannot run (f64 scan, keys-only merge) is refused at
    submit; the requests around it are served."""
    from bulk_amd import batch as B
    from bulk_amd import serve as S
    srv = S.Server(hbm_budget_bytes=1 << 30, max_inflight=2, max_batch_jobs=8,
                   max_batch_elements=1 << 22)
    try:
        a = _ints(100000, seed=5)
        ya = torch.empty_like(a)
        srv.submit(B.scan_job(a, ya), 1 << 20)
        f = torch.rand(5000, dtype=torch.float64, device="cuda")
        yf = torch.empty_like(f)
        with pytest.raises(C.BulkError, match="UNSUPPORTED_DTYPE"):
            srv.submit(B.scan_job(f, yf), 1 << 20)
        # keys-only merge over host buffers: staging accepts it, the batch cannot run it
        k1 = torch.arange(100, dtype=torch.int32).pin_memory()
        k2 = torch.arange(50, dtype=torch.int32).pin_memory()
        ko = torch.empty(150, dtype=torch.int32).pin_memory()
        j = B.Job(C.JOB_MERGE_BY_KEY, C.INT32, -1, 100, 50,
                  {"in0": k1, "in2": k2, "out0": ko})
        with pytest.raises(C.BulkError):
            srv.submit_host(j)
        b = _ints(70000, seed=6)
        yb = torch.empty_like(b)
        srv.submit(B.scan_job(b, yb), 1 << 20)
        srv.drain()
        assert torch.equal(ya.cpu(), torch.cumsum(a.cpu(), 0, dtype=torch.int32))
        assert torch.equal(yb.cpu(), torch.cumsum(b.cpu(), 0, dtype=torch.int32))
        assert srv.stats()["completed"] == 2
    finally:
        srv.close()
