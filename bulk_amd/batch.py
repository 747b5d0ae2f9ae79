"""Python front end of the batch executor (bulkx_batch_* in include/bulkx.h).

A job list is a sequence of ``Job`` records over device tensors; ``Batch``
runs the whole list as segmented work (one launch per op family over the
tiles of all its jobs, plus the radix passes of the larger sort jobs),
recorded into a hipGraph on first use, or as one plain call per job (the
per-call baseline).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import torch

from . import _capi as C
from .api import dtype_code


@dataclass
class Job:
    kind: int
    dtype: int
    value_dtype: int = -1
    n0: int = 0
    n1: int = 0
    tensors: dict = field(default_factory=dict)  # name -> tensor (kept alive)

    def to_c(self) -> C.Job:
        t = self.tensors

        def p(name):
            x = t.get(name)
            return ctypes.c_void_p(x.data_ptr()) if x is not None else ctypes.c_void_p(0)

        return C.Job(self.kind, self.dtype, self.value_dtype, 0, self.n0, self.n1,
                     p("in0"), p("in1"), p("in2"), p("in3"), p("out0"), p("out1"), p("out2"))

    @property
    def elements(self) -> int:
        return self.n0 + (self.n1 if self.kind == C.JOB_MERGE_BY_KEY else 0)


def scan_job(x: torch.Tensor, out: torch.Tensor) -> Job:
    return Job(C.JOB_INCLUSIVE_SCAN, dtype_code(x), -1, x.numel(), 0, {"in0": x, "out0": out})


def reduce_job(x: torch.Tensor, out: torch.Tensor) -> Job:
    return Job(C.JOB_REDUCE, dtype_code(x), -1, x.numel(), 0, {"in0": x, "out0": out})


def reduce_by_key_job(keys, vals, kout, vout, count) -> Job:
    return Job(C.JOB_REDUCE_BY_KEY, dtype_code(keys), dtype_code(vals), keys.numel(), 0,
               {"in0": keys, "in1": vals, "out0": kout, "out1": vout, "out2": count})


def merge_by_key_job(k1, v1, k2, v2, kout, vout) -> Job:
    return Job(C.JOB_MERGE_BY_KEY, dtype_code(k1), dtype_code(v1), k1.numel(), k2.numel(),
               {"in0": k1, "in1": v1, "in2": k2, "in3": v2, "out0": kout, "out1": vout})


def sort_by_key_job(keys, vals) -> Job:
    return Job(C.JOB_SORT_BY_KEY, dtype_code(keys), dtype_code(vals), keys.numel(), 0,
               {"out0": keys, "out1": vals})


class Batch:
    """bulkx_batch: create once with capacities, set a job list, launch."""

    def __init__(self, max_jobs: int, max_elements: int, stream: torch.cuda.Stream | None = None):
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        self._h = ctypes.c_void_p()
        C.check(C.lib().bulkx_batch_create(int(max_jobs), int(max_elements),
                                           ctypes.c_void_p(self.stream.cuda_stream),
                                           ctypes.byref(self._h)), "bulkx_batch_create")
        self._jobs: list[Job] = []

    def set_jobs(self, jobs: list[Job]) -> None:
        arr = (C.Job * max(len(jobs), 1))(*[j.to_c() for j in jobs])
        C.check(C.lib().bulkx_batch_set_jobs(self._h, arr, len(jobs)), "bulkx_batch_set_jobs")
        self._jobs = list(jobs)

    def launch(self, use_graph: bool | int = True) -> None:
        """use_graph True / 1: graph replay; False / 0: eager launches on the
        batch's stream (timed per stage); 2: the graph's branches launched
        eagerly on side streams."""
        C.check(C.lib().bulkx_batch_launch(self._h, int(use_graph)), "bulkx_batch_launch")

    def launch_per_call(self) -> None:
        C.check(C.lib().bulkx_batch_launch_per_call(self._h), "bulkx_batch_launch_per_call")

    def status(self) -> int:
        """Waits for the last launch; BULKX_ERR_LOOKBACK (11) if one of its
        look-back waits gave up, else 0."""
        return int(C.lib().bulkx_batch_status(self._h))

    def reserve(self, elements: int) -> None:
        """Allocate now the scratch job lists of up to `elements` elements need."""
        C.check(C.lib().bulkx_batch_reserve(self._h, int(elements)), "bulkx_batch_reserve")

    def last_timing(self) -> tuple[float, float]:
        """(phase-0 ms, sort-merge-rounds ms) of the last eager launch, from
        HIP events recorded on the batch's stream."""
        a, b = ctypes.c_float(), ctypes.c_float()
        C.check(C.lib().bulkx_batch_last_timing(self._h, ctypes.byref(a), ctypes.byref(b)),
                "bulkx_batch_last_timing")
        return a.value, b.value

    def scratch_bytes(self) -> tuple[int, int]:
        """(sort / reduce_by_key scratch, per-call workspace) bytes the batch holds."""
        t, w = ctypes.c_size_t(), ctypes.c_size_t()
        C.check(C.lib().bulkx_batch_scratch_bytes(self._h, ctypes.byref(t), ctypes.byref(w)),
                "bulkx_batch_scratch_bytes")
        return int(t.value), int(w.value)

    def family_shapes(self) -> dict:
        """{family stage name: (shape, tile elements)} of the first seven
        stages: the tile shape each op family of this batch runs (read from
        the tuning table when the batch was created)."""
        sh = (ctypes.c_int * 7)()
        te = (ctypes.c_int64 * 7)()
        C.check(C.lib().bulkx_batch_family_shapes(self._h, sh, te), "bulkx_batch_family_shapes")
        return {name: (sh[i], te[i]) for i, name in enumerate(self.STAGES[:7])}

    STAGES = ("scan_i32", "scan_f32", "reduce_by_key_i32", "reduce_by_key_f32",
              "reduce_by_key_i64", "merge_by_key", "sort_block", "sort_radix",
              "sort_downsweep_kernels", "sort_upsweep_kernels", "sort_count_scan_kernels")

    @staticmethod
    def sort_geometry() -> tuple[int, int]:
        """(largest sort job block-sorted in one tile, radix tile)."""
        run, tile = ctypes.c_int64(), ctypes.c_int64()
        C.check(C.lib().bulkx_batch_sort_geometry(ctypes.byref(run), ctypes.byref(tile)),
                "bulkx_batch_sort_geometry")
        return run.value, tile.value

    @staticmethod
    def sort_passes(n: int) -> int:
        """Radix passes the batch runs for a sort job of n elements (0: the
        job is one block-sorted tile)."""
        run, _ = Batch.sort_geometry()
        return 0 if n <= run else 4

    def stage_timing(self) -> dict[str, float]:
        """Per-stage ms of the last eager launch (HIP events on the batch's stream)."""
        arr = (ctypes.c_float * len(self.STAGES))()
        C.check(C.lib().bulkx_batch_stage_timing(self._h, arr, len(self.STAGES)),
                "bulkx_batch_stage_timing")
        return dict(zip(self.STAGES, list(arr)))

    def close(self) -> None:
        if self._h:
            C.lib().bulkx_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass
