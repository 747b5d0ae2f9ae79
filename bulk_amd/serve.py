"""Python front end of the serving loop (bulkx_server_* in include/bulkx.h)."""
from __future__ import annotations

import ctypes

from . import _capi as C
from .batch import Job

SERVE_PER_CALL = 1
SERVE_NO_GRAPH = 2


class Server:
    """Continuous-batching request loop over device-resident jobs."""

    def __init__(self, hbm_budget_bytes: int, max_inflight: int = 2, max_batch_jobs: int = 512,
                 max_batch_elements: int = 1 << 26, flags: int = 0):
        self._h = ctypes.c_void_p()
        C.check(C.lib().bulkx_server_create(int(hbm_budget_bytes), int(max_inflight),
                                            int(max_batch_jobs), int(max_batch_elements),
                                            int(flags), ctypes.byref(self._h)),
                "bulkx_server_create")
        self._keep: list[Job] = []

    def now_us(self) -> float:
        return float(C.lib().bulkx_server_now_us(self._h))

    def submit(self, job: Job, nbytes: int, arrival_us: float | None = None) -> int:
        rid = ctypes.c_int64()
        cj = job.to_c()
        if arrival_us is None:
            st = C.lib().bulkx_server_submit(self._h, ctypes.byref(cj), int(nbytes), ctypes.byref(rid))
        else:
            st = C.lib().bulkx_server_submit_at(self._h, ctypes.byref(cj), int(nbytes),
                                                float(arrival_us), ctypes.byref(rid))
        C.check(st, "bulkx_server_submit")
        self._keep.append(job)
        return rid.value

    def submit_host(self, job: Job, arrival_us: float | None = None) -> tuple[int, int]:
        """Submit a job over host tensors (pinned for asynchronous copies):
        staged to the device when admitted, results copied back.  Returns
        (request id, device bytes charged to the budget)."""
        rid, nbytes = ctypes.c_int64(), ctypes.c_int64()
        cj = job.to_c()
        C.check(C.lib().bulkx_server_submit_host(self._h, ctypes.byref(cj),
                                                 -1.0 if arrival_us is None else float(arrival_us),
                                                 ctypes.byref(rid), ctypes.byref(nbytes)),
                "bulkx_server_submit_host")
        self._keep.append(job)
        return rid.value, nbytes.value

    def drain(self) -> None:
        C.check(C.lib().bulkx_server_drain(self._h), "bulkx_server_drain")

    def times(self, rid: int) -> tuple[float, float, float]:
        a, l, d = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        C.check(C.lib().bulkx_server_request_times(self._h, rid, ctypes.byref(a), ctypes.byref(l),
                                                   ctypes.byref(d)), "bulkx_server_request_times")
        return a.value, l.value, d.value

    def stats(self) -> dict:
        b, c, p = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        C.check(C.lib().bulkx_server_stats(self._h, ctypes.byref(b), ctypes.byref(c), ctypes.byref(p)),
                "bulkx_server_stats")
        return {"batches": b.value, "completed": c.value, "peak_bytes": p.value}

    def close(self) -> None:
        if self._h:
            C.lib().bulkx_server_destroy(self._h)
            self._h = ctypes.c_void_p()
        self._keep = []

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass
