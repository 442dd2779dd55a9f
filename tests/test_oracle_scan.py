"""The restated Bolt scan-by-key (kernel/boltScan.cl) against brute force."""
import numpy as np
import pytest

from oracle import pyoracle as O


def brute(keys, vals):
    out = np.empty(len(keys), dtype=np.int64)
    run, prev = 0, None
    for i, (k, v) in enumerate(zip(keys, vals)):
        run = v if k != prev else run + v
        prev = k
        out[i] = run - 1
    return out


@pytest.mark.parametrize("n", [1, 2, 127, 128, 129, 1000, 16385, 40000, 300001])
def test_bolt_matches_brute_force(n):
    rng = np.random.default_rng(n)
    cuts = np.sort(rng.integers(0, n + 1, 2))
    keys = np.ones(n, dtype=np.uint32) + (np.arange(n) >= cuts[0]) + (np.arange(n) >= cuts[1])
    vals = rng.integers(0, 2, n).astype(np.int32)
    got = O.bolt_scan(keys, vals)
    assert np.array_equal(got.astype(np.int64), brute(keys, vals))


def test_bolt_single_segment_and_all_ones():
    n = 128 * 128 * 3 + 17  # several work-items per thread in the middle kernel
    keys = np.full(n, 2, dtype=np.uint32)
    vals = np.ones(n, dtype=np.int32)
    assert np.array_equal(O.bolt_scan(keys, vals), np.arange(n, dtype=np.int32))
