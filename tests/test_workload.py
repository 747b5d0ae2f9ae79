"""Host-side checks of the mixed-job workload description (no GPU)."""
import numpy as np

from bulk_amd import _capi as C
from bulk_amd import workload as W


def test_mix_sizes_and_dtypes():
    specs = W.make_specs(4096, seed=5)
    kinds = np.bincount([s.kind for s in specs], minlength=4)
    for kind, p in W.MIX:
        assert abs(kinds[kind] / 4096 - p) < 0.03
    n = np.array([s.n for s in specs])
    assert n.min() >= 1 << 10 and n.max() <= 1 << 22
    # log-uniform: the mean of 4096 draws is near (2^22 - 2^10) / (12 ln 2)
    assert abs(n.mean() / W.mean_log_uniform() - 1) < 0.08
    assert abs(np.mean([s.dtype == "int32" for s in specs]) - 0.5) < 0.03
    # the stream fits int32 element offsets and the batch tile table
    assert n.sum() < 2 ** 31


def test_specs_are_deterministic():
    a = W.make_specs(64, seed=9)
    b = W.make_specs(64, seed=9)
    assert a == b
    assert a != W.make_specs(64, seed=10)


def test_poisson_arrivals_rate():
    arr = W.poisson_arrivals(20000, rate_per_s=1e5, seed=1)
    assert np.all(np.diff(arr) >= 0)
    # mean gap 10 us within 3%
    assert abs(arr[-1] / 20000 - 10.0) < 0.3
