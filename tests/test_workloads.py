"""Scenario sets of the batched and serving workloads, and the autotuner's
size buckets (CPU)."""
import math

import numpy as np

from sampo_amd import batch as B
from sampo_amd import workloads as W


def test_ensemble_set():
    s = W.ensemble()
    assert len(s) == 1024
    n0 = np.array([x.n0 for x in s])
    assert n0.min() >= 100_000 and n0.max() <= 1_000_000
    assert [x.seed for x in s] == list(range(1000, 2024))
    assert all(x.days == 365 and x.dmr == W.DMR for x in s)
    assert W.ensemble()[17].n0 == s[17].n0  # a pure function of its arguments


def test_sweep_set_ranges():
    s = W.sweep()
    assert len(s) == 4096
    n0 = np.array([x.n0 for x in s])
    assert n0.min() >= 1000 and n0.max() <= 4_000_000
    # log-uniform: the median sits near the geometric mean of the range
    assert abs(math.log(np.median(n0)) - 0.5 * (math.log(1e3) + math.log(4e6))) < 0.2
    assert all(18.0 <= x.t_mean <= 32.0 and 2.0 <= x.t_amp <= 8.0 for x in s)
    for x in s[:200]:
        for c in range(4):
            assert 0.5 * W.DMR[c] - 1e-12 <= x.dmr[c] <= 1.5 * W.DMR[c] + 1e-12
    assert [x.seed for x in s[:3]] == [100000, 100001, 100002]


def test_serving_arrivals():
    s = W.serving(rate=50.0, duration=20.0)
    t = np.array([x.arrival for x in s])
    assert np.all(np.diff(t) > 0) and t[-1] < 20.0
    assert 800 < len(s) < 1200  # Poisson(1000)
    assert all(30 <= x.days <= 365 for x in s)
    n0 = np.array([x.n0 for x in s])
    assert n0.min() >= 1e4 and n0.max() <= 1e7


def test_size_buckets():
    assert B.bucket_of(1000) == 0 and B.bucket_of(3999) == 0 and B.bucket_of(4000) == 1
    assert B.bucket_of(4_000_000) == 5 and B.bucket_of(5_000_000) == 6 and B.bucket_of(10) == 0
    for b in range(7):
        lo, hi = B.bucket_range(b)
        assert B.bucket_of(int(lo) + 1) == b and B.bucket_of(int(hi) - 1) == b
    t = B.Tuning(chunk={0: 4096, 5: 65536})
    assert t.chunk_for(1500) == 4096 and t.chunk_for(2_000_000) == 65536 and t.chunk_for(50_000) == B.DEFAULT_CHUNK
