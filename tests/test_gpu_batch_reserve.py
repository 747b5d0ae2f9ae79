"""bulkx_batch_reserve: after reserving for E elements, every job list of at
most E elements (and at most max_jobs jobs) is set without growing the
batch's scratch.  The job lists are the worst cases of the per-job terms:
many reduce_by_key jobs of int64 values one element past a tile (two tiles,
2 * (T + 1) run slots each), many sort jobs one key past the block-sorted
size (the radix passes' ping-pong space and a partial tile's counts), both
mixed, one job of all E elements of each kind, and the reduce_by_key
families' second tile shape.  Nothing is launched: set_jobs only lays out and
uploads the job table, so the jobs may share one buffer."""
import pytest
import torch

from bulk_amd import _capi as C

pytestmark = pytest.mark.gpu

E = 1 << 22
MAX_JOBS = 1024


def _rbk_jobs(B, buf32, buf64, cnt, n, count):
    return [B.reduce_by_key_job(buf32[:n], buf64[:n], buf32[:n], buf64[:n], cnt) for _ in range(count)]


def _sort_jobs(B, buf32, n, count):
    return [B.sort_by_key_job(buf32[:n], buf32[:n]) for _ in range(count)]


def _check_no_growth(B, jobs):
    assert sum(j.elements for j in jobs) <= E and len(jobs) <= MAX_JOBS
    b = B.Batch(max_jobs=MAX_JOBS, max_elements=E)
    b.reserve(E)
    reserved = b.scratch_bytes()
    b.set_jobs(jobs)
    assert b.scratch_bytes() == reserved, (reserved, b.scratch_bytes())
    b.close()


@pytest.mark.parametrize("case", ["rbk_i64_two_tiles", "sort_radix_min", "mixed", "rbk_one_job",
                                  "sort_one_job"])
def test_reserve_covers_job_lists(bx, case):
    from bulk_amd import batch as B
    buf32 = torch.zeros(E, dtype=torch.int32, device="cuda")
    buf64 = torch.zeros(E, dtype=torch.int64, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    run, _ = B.Batch.sort_geometry()
    rbk_tile = 8192  # the reduce_by_key families' default tile
    if case == "rbk_i64_two_tiles":
        jobs = _rbk_jobs(B, buf32, buf64, cnt, rbk_tile + 1, E // (rbk_tile + 1))
    elif case == "sort_radix_min":
        jobs = _sort_jobs(B, buf32, run + 1, E // (run + 1))
    elif case == "mixed":
        half = E // 2
        jobs = (_rbk_jobs(B, buf32, buf64, cnt, rbk_tile + 1, half // (rbk_tile + 1)) +
                _sort_jobs(B, buf32, run + 1, half // (run + 1)))
    elif case == "rbk_one_job":
        jobs = _rbk_jobs(B, buf32, buf64, cnt, E, 1)
    else:
        jobs = _sort_jobs(B, buf32, E, 1)
    print(f"reserve case={case} jobs={len(jobs)}", flush=True)
    _check_no_growth(B, jobs)


def test_reserve_covers_other_rbk_shape(bx):
    """The reduce_by_key families' second tile shape (set through the tuning
    table before the batch is created), jobs one element past its tile."""
    from bulk_amd import batch as B
    lib = C.lib()
    buf32 = torch.zeros(E, dtype=torch.int32, device="cuda")
    buf64 = torch.zeros(E, dtype=torch.int64, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    try:
        C.check(lib.bulkx_set_tuning(C.ALGO_BATCH_REDUCE_BY_KEY, C.INT64, 0, 1), "bulkx_set_tuning")
        b = B.Batch(max_jobs=MAX_JOBS, max_elements=E)
        shape, tile = b.family_shapes()["reduce_by_key_i64"]
        assert shape == 1
        b.close()
        jobs = _rbk_jobs(B, buf32, buf64, cnt, tile + 1, min(MAX_JOBS, E // (tile + 1)))
        print(f"reserve case=rbk_i64_shape1 tile={tile} jobs={len(jobs)}", flush=True)
        _check_no_growth(B, jobs)
    finally:
        C.check(lib.bulkx_load_tuning(None, None), "bulkx_load_tuning")
