"""The engine's arithmetic shortcuts, checked as identities (CPU).

Where the fused agent step computes a reference expression by cheaper means,
the two must agree bit for bit on every input the engine admits.  Each
identity is checked here exhaustively, or over the special values and a dense
random sample of float bit patterns, with numpy's IEEE arithmetic standing in
for the device's (both round to nearest even and keep denormals).
"""
import numpy as np

from oracle import pyoracle as O

F32 = np.float32


def finite_floats(n, seed):
    rng = np.random.default_rng(seed)
    b = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    x = b.view(F32)
    return x[np.isfinite(x)]


def test_age_in_days_is_integer_division():
    """killAgents.cl:55-147 and calcL1de take (uint)min(age / 24.0f, 99.0f)
    and (uint)(age / 24.0f) of an age in whole hours; admitted ages are
    integers below 2^16, for which both are integer divisions."""
    age = np.arange(0, 1 << 16, dtype=np.uint32)
    q = age.astype(F32) / F32(24.0)
    assert np.array_equal(np.minimum(q, F32(99.0)).astype(np.uint32), np.minimum(age // 24, 99))
    assert np.array_equal(q.astype(np.uint32), age // 24)


def development_increment(t):
    """updateAgents.cl:69-70's increment, 1/((111/(T-16))*14): the quotient in
    binary32, the rest in binary64."""
    with np.errstate(divide="ignore", over="ignore"):
        q = F32(111.0) / (t - F32(16.0))
        return 1.0 / (q.astype(np.float64) * 14.0)


def test_zero_increment_has_the_sign_of_t_minus_16():
    """Without the condition the reference adds 0 * increment; for finite T
    that zero's sign is the sign of T - 16 (the increment is +0 at 16)."""
    t = np.concatenate([
        finite_floats(1 << 22, 1),
        np.array([16.0, -0.0, 0.0, 15.999999, 16.000002, np.finfo(F32).max, -np.finfo(F32).max,
                  np.finfo(F32).tiny, -np.finfo(F32).tiny], F32),
        np.nextafter(F32(16.0), F32(0.0)) - np.arange(0, 4096, dtype=F32) * F32(2.0**-19),
        F32(16.0) + np.arange(0, 4096, dtype=F32) * F32(2.0**-19),
        (np.arange(-200000, 200000, dtype=np.int64) * 1e-3).astype(F32),
    ])
    inc = development_increment(t)
    assert np.isfinite(inc).all()
    with np.errstate(invalid="ignore"):
        z = 0.0 * inc
    assert np.array_equal(np.signbit(z), np.signbit(t - F32(16.0)))


def test_adding_a_zero_in_binary64_is_adding_it_in_binary32():
    """(float)((double)csd + z) == csd + (float)z, as bits, for z = +0 and -0
    and every class of csd: zeros, denormals, normals, infinities, NaNs."""
    x = np.concatenate([
        np.random.default_rng(2).integers(0, 2**32, 1 << 22, dtype=np.uint64).astype(np.uint32).view(F32),
        np.array([0.0, -0.0, np.inf, -np.inf, np.nan], F32),
        np.array([1, 0x80000001, 0x007fffff, 0x807fffff, 0x7fc00001, 0xffc00001], np.uint32).view(F32),
    ])
    for z in (0.0, -0.0):
        with np.errstate(invalid="ignore"):
            wide = (x.astype(np.float64) + z).astype(F32)
            narrow = x + F32(z)
        assert np.array_equal(wide.view(np.uint32), narrow.view(np.uint32))


def taus_step(z, s1, s2, s3, m):
    b = ((z << np.uint32(s1)) ^ z) >> np.uint32(s2)
    return ((z & np.uint32(m)) << np.uint32(s3)) ^ b


def test_prepared_seed_sets_draw_what_hybrid_taus_draws():
    """The engine prepares each seed set once (the first and third Tausworthe
    components xored, z2 + 128, 1664525 * z4) and draws from it with one
    step, one multiply-add and two xors; the draws equal the oracle's
    HybridTaus on random seed sets and indices, the extreme ones included."""
    rng = np.random.default_rng(4)
    n = 4096
    z = rng.integers(0, 2**32, (n, 4), dtype=np.uint64).astype(np.uint32)
    z[:8] = [[0, 0, 0, 0], [2**32 - 1] * 4, [2**32 - 128] * 4, [127, 2**32 - 129, 128, 1],
             [1, 2, 3, 4], [2**31, 2**31, 2**31, 2**31], [5, 2**32 - 128, 7, 2**31 - 1], [9, 9, 9, 9]]
    gid = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    gid[:8] = [0, 1, 2**31, 2**32 - 1, 64, 2**32 - 128, 12345, 2**31 - 1]
    with np.errstate(over="ignore"):
        k = taus_step(z[:, 0] + np.uint32(128), 13, 19, 12, 0xfffffffe) ^ \
            taus_step(z[:, 2] + np.uint32(128), 3, 11, 17, 0xfffffff0)
        b = z[:, 1] + np.uint32(128)
        m = np.uint32(1664525) * z[:, 3]
        bits = k ^ taus_step(b + gid, 2, 25, 4, 0xfffffff8) ^ (m * gid + np.uint32(1013904223))
    draws = F32(2.0**-32) * bits.astype(F32)
    want = np.array([O.hybrid_taus(z[i], int(gid[i])) for i in range(n)], F32)
    assert np.array_equal(draws.view(np.uint32), want.view(np.uint32))
