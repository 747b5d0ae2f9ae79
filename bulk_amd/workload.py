"""Synthetic mixed-job stream: 4096 uneven scan / reduce_by_key / merge /
sort_by_key jobs with Poisson arrivals, generated device-resident.

Job shapes are drawn on the host from a numpy Generator (so the list is the
same on every machine); the data of each job is generated on the GPU with a
seeded torch Generator, before any timing starts.

Per-job distributions follow the single-op workloads:
  * scan      -- int32 uniform [-1000, 1000]; float32 uniform [0,1) or
                 normal(0,1) (alternating)
  * reduce_by_key -- int32 keys in contiguous runs, run length geometric with
                 mean 8, key = run index; values int32 uniform [-2^20, 2^20] or
                 float32 uniform [-1, 1]
  * merge     -- two sorted key arrays (sizes n//2 and n - n//2), 10% of the
                 entries drawn from a shared 1024-value pool to force ties;
                 float32 uniform [0,1) or int32 uniform [-2^30, 2^30);
                 payload int32 (first range i, second range 2^28 + i)
  * sort_by_key -- int32 keys over the full range or float32 uniform [0,1);
                 payload = original index
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _capi as C
from . import batch as B

MIX = ((C.JOB_INCLUSIVE_SCAN, 0.4), (C.JOB_REDUCE_BY_KEY, 0.3), (C.JOB_MERGE_BY_KEY, 0.2),
       (C.JOB_SORT_BY_KEY, 0.1))
KIND_NAMES = {C.JOB_INCLUSIVE_SCAN: "scan", C.JOB_REDUCE_BY_KEY: "reduce_by_key",
              C.JOB_MERGE_BY_KEY: "merge_by_key", C.JOB_SORT_BY_KEY: "sort_by_key",
              C.JOB_REDUCE: "reduce"}
MERGE_PAYLOAD_B = 1 << 28


@dataclass
class Spec:
    kind: int
    dtype: str  # "int32" or "float32"
    n: int
    variant: int  # distribution variant (scan: 0 uniform / 1 normal)


def make_specs(num_jobs: int = 4096, seed: int = 5, log2_min: int = 10, log2_max: int = 22,
               mix=MIX) -> list[Spec]:
    """Job kinds by the mix, sizes log-uniform on [2^log2_min, 2^log2_max],
    dtype int32 / float32 with probability 1/2 each."""
    rng = np.random.default_rng(seed)
    kinds = np.array([k for k, _ in mix])
    probs = np.array([p for _, p in mix], dtype=np.float64)
    probs /= probs.sum()
    ks = rng.choice(kinds, size=num_jobs, p=probs)
    u = rng.random(num_jobs)
    sizes = np.floor(np.exp2(log2_min + (log2_max - log2_min) * u)).astype(np.int64)
    dts = rng.integers(0, 2, size=num_jobs)
    var = rng.integers(0, 2, size=num_jobs)
    return [Spec(int(k), "int32" if d == 0 else "float32", int(n), int(v))
            for k, n, d, v in zip(ks, sizes, dts, var)]


def poisson_arrivals(num_jobs: int, rate_per_s: float, seed: int = 6) -> np.ndarray:
    """Arrival offsets in microseconds (exponential inter-arrival gaps)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1e6 / rate_per_s, size=num_jobs)
    return np.cumsum(gaps)


def _runs_keys(n: int, mean: float, g: torch.Generator, dev) -> torch.Tensor:
    m = max(1, int(n / mean * 1.25) + 64)
    while True:
        lens = torch.empty(m, dtype=torch.float32, device=dev).geometric_(1.0 / mean, generator=g)
        lens = lens.to(torch.int64)
        if int(lens.sum().item()) >= n:
            break
        m *= 2
    keys = torch.repeat_interleave(torch.arange(m, dtype=torch.int32, device=dev), lens)
    return keys[:n].contiguous()


def _merge_keys(n: int, dtype: str, pool: torch.Tensor, g: torch.Generator, dev) -> torch.Tensor:
    if dtype == "float32":
        k = torch.rand(n, generator=g, device=dev, dtype=torch.float32)
    else:
        k = torch.randint(-(1 << 30), 1 << 30, (n,), generator=g, device=dev, dtype=torch.int32)
    sel = torch.rand(n, generator=g, device=dev) < 0.1
    idx = torch.randint(0, pool.numel(), (n,), generator=g, device=dev)
    k = torch.where(sel, pool[idx], k)
    return torch.sort(k, stable=True).values.contiguous()


class JobData:
    """Device tensors of one job plus its batch record."""

    def __init__(self, spec: Spec, job: B.Job, restore=None):
        self.spec = spec
        self.job = job
        self.restore = restore  # (keys, vals, keys0, vals0) for in-place sorts

    @property
    def elements(self) -> int:
        return self.spec.n

    def footprint_bytes(self) -> int:
        """Device bytes the job's inputs, outputs and scratch occupy."""
        n = self.spec.n
        if self.spec.kind == C.JOB_INCLUSIVE_SCAN:
            return 8 * n
        if self.spec.kind == C.JOB_REDUCE_BY_KEY:
            return 16 * n + 8
        if self.spec.kind == C.JOB_MERGE_BY_KEY:
            return 16 * n
        return 16 * n  # sort: keys + payload in place, plus ping-pong space

    def algorithmic_bytes(self, runs: int | None = None) -> int:
        """Compulsory HBM bytes (each input read once, each output written
        once).  reduce_by_key needs the run count."""
        n = self.spec.n
        k = self.spec.kind
        if k == C.JOB_INCLUSIVE_SCAN:
            return 2 * n * 4
        if k == C.JOB_REDUCE_BY_KEY:
            return n * 8 + (runs if runs is not None else 0) * 8 + 8
        return 2 * n * 8  # merge (n = n1 + n2) and sort: keys + payload in and out

    def reset(self) -> None:
        if self.restore is not None:
            k, v, k0, v0 = self.restore
            k.copy_(k0)
            v.copy_(v0)


def _tdt(dtype: str):
    return torch.int32 if dtype == "int32" else torch.float32


def build_jobs(specs: list[Spec], seed: int = 7, device="cuda") -> list[JobData]:
    dev = torch.device(device)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    pool_f = torch.rand(1024, generator=g, device=dev, dtype=torch.float32)
    pool_i = torch.randint(-(1 << 30), 1 << 30, (1024,), generator=g, device=dev, dtype=torch.int32)
    out = []
    for s in specs:
        n, dt = s.n, _tdt(s.dtype)
        if s.kind == C.JOB_INCLUSIVE_SCAN:
            if s.dtype == "int32":
                x = torch.randint(-1000, 1001, (n,), generator=g, device=dev, dtype=torch.int32)
            elif s.variant == 0:
                x = torch.rand(n, generator=g, device=dev, dtype=torch.float32)
            else:
                x = torch.randn(n, generator=g, device=dev, dtype=torch.float32)
            y = torch.empty_like(x)
            out.append(JobData(s, B.scan_job(x, y)))
        elif s.kind == C.JOB_REDUCE_BY_KEY:
            keys = _runs_keys(n, 8.0, g, dev)
            if s.dtype == "int32":
                vals = torch.randint(-(1 << 20), (1 << 20) + 1, (n,), generator=g, device=dev,
                                     dtype=torch.int32)
            else:
                vals = torch.rand(n, generator=g, device=dev, dtype=torch.float32) * 2 - 1
            ko = torch.empty(n, dtype=torch.int32, device=dev)
            vo = torch.empty(n, dtype=dt, device=dev)
            cnt = torch.zeros(1, dtype=torch.int64, device=dev)
            out.append(JobData(s, B.reduce_by_key_job(keys, vals, ko, vo, cnt)))
        elif s.kind == C.JOB_MERGE_BY_KEY:
            n1 = n // 2
            n2 = n - n1
            pool = pool_f if s.dtype == "float32" else pool_i
            a = _merge_keys(n1, s.dtype, pool, g, dev)
            b = _merge_keys(n2, s.dtype, pool, g, dev)
            va = torch.arange(n1, dtype=torch.int32, device=dev)
            vb = MERGE_PAYLOAD_B + torch.arange(n2, dtype=torch.int32, device=dev)
            ko = torch.empty(n, dtype=dt, device=dev)
            vo = torch.empty(n, dtype=torch.int32, device=dev)
            out.append(JobData(s, B.merge_by_key_job(a, va, b, vb, ko, vo)))
        else:
            if s.dtype == "int32":
                keys = torch.randint(-(1 << 31), (1 << 31) - 1, (n,), generator=g, device=dev,
                                     dtype=torch.int32)
            else:
                keys = torch.rand(n, generator=g, device=dev, dtype=torch.float32)
            vals = torch.arange(n, dtype=torch.int32, device=dev)
            keys0, vals0 = keys.clone(), vals.clone()
            out.append(JobData(s, B.sort_by_key_job(keys, vals), (keys, vals, keys0, vals0)))
    return out


def reset_all(jobs: list[JobData]) -> None:
    for j in jobs:
        j.reset()


def total_elements(jobs) -> int:
    return sum(j.elements for j in jobs)


def run_counts(jobs) -> dict[int, int]:
    """Run counts of the reduce_by_key jobs (after they ran)."""
    return {i: int(j.job.tensors["out2"].item()) for i, j in enumerate(jobs)
            if j.spec.kind == C.JOB_REDUCE_BY_KEY}


def mean_log_uniform(log2_min: int = 10, log2_max: int = 22) -> float:
    return (2.0 ** log2_max - 2.0 ** log2_min) / ((log2_max - log2_min) * math.log(2.0))
