"""Pins for the oracle's numerics and random numbers (CPU).

Each test checks an oracle function against something independent of its
code: the host libm/numpy, the host glibc, mathematical properties (periods,
continuity, closed forms), or constants read from the reference sources
(tests/golden/constants.json, made by tests/golden/make_constants.py).
"""
import ctypes
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CONST = json.load(open(os.path.join(ROOT, "tests", "golden", "constants.json")))


def ulps(a, b):
    a = np.asarray(a, dtype=np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, dtype=np.float32).view(np.int32).astype(np.int64)
    a = np.where(a < 0, np.int64(-2**31) - a, a)
    b = np.where(b < 0, np.int64(-2**31) - b, b)
    return np.abs(a - b)


def sample_floats(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(lo, hi, n).astype(np.float32)


def test_exp_log_sin_pow_within_one_ulp_of_libm():
    L = O.lib()
    x = sample_floats(20000, -87.0, 88.0, 1)
    e = np.array([L.or_expf(float(v)) for v in x], dtype=np.float32)
    assert ulps(e, np.exp(x.astype(np.float64)).astype(np.float32)).max() <= 1
    x = np.concatenate([sample_floats(10000, 1e-30, 1.0, 2), sample_floats(10000, 1.0, 1e30, 3),
                        np.float32([2.0**-149, 1.0, 2.0, 0.5])])
    lg = np.array([L.or_logf(float(v)) for v in x], dtype=np.float32)
    assert ulps(lg, np.log(x.astype(np.float64)).astype(np.float32)).max() <= 1
    x = sample_floats(20000, 0.0, 2 * np.pi, 4)
    sn = np.array([L.or_sinf(float(v)) for v in x], dtype=np.float32)
    assert ulps(sn, np.sin(x.astype(np.float64)).astype(np.float32)).max() <= 1
    xs, ys = sample_floats(20000, 1e-3, 1.0, 5), sample_floats(20000, -30.0, 30.0, 6)
    pw = np.array([L.or_powf(float(a), float(b)) for a, b in zip(xs, ys)], dtype=np.float32)
    with np.errstate(over="ignore"):
        ref = np.power(xs.astype(np.float64), ys.astype(np.float64)).astype(np.float32)
    assert ulps(pw, ref).max() <= 1


def test_special_values():
    L = O.lib()
    assert L.or_logf(0.0) == -np.inf and np.isnan(L.or_logf(-1.0))
    assert L.or_expf(0.0) == 1.0 and L.or_expf(-200.0) == 0.0 and L.or_expf(100.0) == np.inf
    assert L.or_powf(0.8, 0.0) == 1.0 and L.or_powf(1.0, 7.5) == 1.0
    assert L.or_sinf(0.0) == 0.0


def test_sqrt_is_correctly_rounded():
    L = O.lib()
    x = np.concatenate([sample_floats(20000, 0.0, 1e6, 7), sample_floats(2000, 0.0, 1e-38, 8)])
    got = np.array([L.or_sqrtf(float(v)) for v in x], dtype=np.float32)
    assert np.array_equal(got, np.sqrt(x))  # numpy float32 sqrt is IEEE


def test_float_to_uint_saturates():
    L = O.lib()
    assert L.or_f2u(float("nan")) == 0
    assert L.or_f2u(-3.0) == 0 and L.or_f2u(-0.99) == 0
    assert L.or_f2u(5e9) == 0xFFFFFFFF and L.or_f2u(float("inf")) == 0xFFFFFFFF
    assert L.or_f2u(4294967040.0) == 4294967040 and L.or_f2u(7.99) == 7


def test_seed_stream_is_glibc_rand():
    libc = ctypes.CDLL("libc.so.6")
    for seed in (1, 2, 12345, 2**31 - 1, 2**31, 4294967295):
        libc.srand(ctypes.c_uint(seed))
        ref = np.array([libc.rand() for _ in range(2000)], dtype=np.uint32)
        assert np.array_equal(O.seed_stream(seed, 2000), ref)


def test_rng_scale_literal_is_two_to_minus_32():
    assert np.float32(float(CONST["rngScale"]["literal"])) == np.float32(2.0**-32)


def test_rng_structure():
    # z2 and gid enter only as z2 + gid; z4 enters only as z4 * gid
    z = np.array([123456789, 362436069, 521288629, 88675123], dtype=np.uint32)
    for gid in (0, 1, 77, 2**31, 2**32 - 1):
        a = O.hybrid_taus([z[0], z[1] + 5, z[2], 0], gid)
        b = O.hybrid_taus([z[0], z[1], z[2], 0], (gid + 5) % 2**32)
        assert a == b
    assert O.hybrid_taus([1, 2, 3, 0], 10) == O.hybrid_taus([1, 2 + 10, 3, 7], 0)
    # outputs are multiples of 2^-32 in [0, 1]
    u = np.array([O.hybrid_taus(z, g) for g in range(20000)], dtype=np.float64)
    assert u.min() >= 0.0 and u.max() <= 1.0
    assert abs(u.mean() - 0.5) < 0.01 and abs(u.var() - 1 / 12) < 0.005
    # streams k = 0..3 of one iteration use independent seed sets: no cross-correlation
    z2 = np.array([987654321, 1234567, 7654321, 55555], dtype=np.uint32)
    v = np.array([O.hybrid_taus(z2, g) for g in range(20000)], dtype=np.float64)
    assert abs(np.corrcoef(u, v)[0, 1]) < 0.03


def test_tausworthe_component_periods():
    """Each component of the generator is a maximal-period Tausworthe
    recurrence (L'Ecuyer 1996): a mask keeping k bits gives period 2^k - 1,
    here 2^31-1, 2^29-1, 2^28-1 for the constants the reference passes
    (include/gpuRand.h:46-48; its comment's "2^30-1" for the second component
    does not match its own mask 4294967288, which keeps 29 bits)."""
    src = r'''
#include <stdio.h>
#include "sampo_canon.h"
int main(void){
  const int p[3][4]={{13,19,12,0},{2,25,4,0},{3,11,17,0}};
  const unsigned m[3]={0xfffffffeu,0xfffffff8u,0xfffffff0u};
  for(int c=0;c<3;c++){ unsigned z0=sc_taus_step(1234567u,p[c][0],p[c][1],p[c][2],m[c]); unsigned z=z0; unsigned long long n=0;
    do { z=sc_taus_step(z,p[c][0],p[c][1],p[c][2],m[c]); n++; } while(z!=z0 && n<(1ull<<33));
    printf("%llu\n", n); }
  return 0; }
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "p.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "p")
        subprocess.run(["gcc", "-O2", "-I", os.path.join(ROOT, "sampo_amd", "csrc"), "-o", exe, c], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    assert [int(v) for v in out] == [2**31 - 1, 2**29 - 1, 2**28 - 1]
    # and those are the constants the reference uses
    assert CONST["taus"] == [[1, 13, 19, 12, 4294967294], [2, 2, 25, 4, 4294967288], [3, 3, 11, 17, 4294967280]]
    a, c_ = CONST["lcg"]
    assert a % 4 == 1 and c_ % 2 == 1  # Hull-Dobell: full period 2^32


def test_box_muller():
    L = O.lib()
    rng = np.random.default_rng(9)
    for u0, u1 in rng.uniform(1e-6, 1.0, (2000, 2)).astype(np.float32):
        got = L.or_box_muller(float(u0), float(u1), 170.0, 30.0)
        ref = np.sqrt(-2 * np.log(np.float64(u0))) * np.sin(2 * np.pi * np.float64(u1)) * 30 + 170
        assert abs(got - ref) < 1e-3 * max(1.0, abs(ref))
    assert L.or_box_muller(1.0, 0.3, 1.0, 0.1) == np.float32(1.0)  # u0 = 1: r = 0


def test_eggs_hatching_matches_reference_coefficients_and_is_continuous():
    L = O.lib()
    h = CONST["eggsHatching"]
    edges = [0.0] + h["thresholds"] + [1.0]
    for i, (m, b) in enumerate(zip(h["m"], h["b"])):
        for p in np.linspace(edges[i], edges[i + 1], 7)[1:-1].astype(np.float32):
            v = np.float32(np.float32(np.float32(m) * p) + np.float32(b)) + np.float32(0.5)
            assert L.or_eggs_hatching(float(p)) == np.float32(max(0, int(v)))
    for t in h["thresholds"]:  # pieces meet at the thresholds
        lo = L.or_eggs_hatching(float(np.nextafter(np.float32(t), np.float32(0))))
        hi = L.or_eggs_hatching(float(t))
        assert abs(hi - lo) <= 1.0


def test_dev_rates_table_matches_reference():
    L = O.lib()
    vals = np.float32(CONST["devRates"]["values"])
    assert len(vals) == 32
    got = np.float32([L.or_larva_development(float(t)) for t in range(12, 44)])
    assert np.array_equal(got, vals)
    assert L.or_larva_development(5.0) == vals[0]
    assert L.or_larva_development(47.0) == vals[31]  # clamped index (reference reads past the table)
    assert L.or_larva_development(30.7) == vals[18]


def test_pupa_delay_is_binary64():
    L = O.lib()
    for t in np.float32([18.0, 25.3, 30.0, 33.7]):
        assert L.or_pupa_delay(float(t)) == np.float32(-0.923 * np.float64(t) + 60.923)


def test_mortality_closed_forms():
    L = O.lib()
    egg = L.or_hmr_reference(0, 5, 0, 160000.0)
    assert abs(egg - (1 - 0.9 ** (1 / 24))) < 2e-7
    assert L.or_hmr_reference(2, 5, 0, 160000.0) == egg  # pupae use the egg rate
    assert L.or_hmr_reference(3, 0, 0, 160000.0) == egg  # adult DMR(0) = a = 0.1 exactly
    a, B, s = CONST["adultMortality"]["a"], CONST["adultMortality"]["B"], CONST["adultMortality"]["s"]
    for d in (1, 10, 50, 99):
        e = np.exp(d / B)
        dmr = a * e / (1 + a * B * s * (e - 1))
        assert abs(L.or_hmr_reference(3, d, 0, 160000.0) - (1 - (1 - dmr) ** (1 / 24))) < 1e-6
    assert L.or_hmr_reference(1, 3, 0, 160000.0) == egg  # no larvae: DMR = 0.1
    cap = 1 - 0.2 ** (1 / 24)
    assert abs(L.or_hmr_reference(1, 1, 10**9, 100.0) - cap) < 1e-6  # capped at 0.8
    dmr = 0.1 * np.exp(50000 / (4 * 160000.0))
    assert abs(L.or_hmr_reference(1, 4, 50000, 160000.0) - (1 - (1 - dmr) ** (1 / 24))) < 1e-6


def test_hybrid_taus_known_answers():
    """HybridTaus bit for bit against tests/kat_rng.py, a numpy restatement
    built only from the constants and state formation parsed out of
    include/gpuRand.h:14-50: a wrong shift, mask, multiplier, offset, or a
    gid entering the wrong component fails."""
    from tests import kat_rng as K
    assert [d["z"] for d in CONST["hybridState"]] == [1, 2, 3, 4]
    rng = np.random.default_rng(20240611)
    special = np.array([0, 1, 2, 3, 127, 128, 2**31 - 1, 2**31, 2**32 - 2, 2**32 - 1], dtype=np.uint64)
    n_checked = 0
    for trial in range(40):
        seeds = rng.integers(0, 2**32, size=4, dtype=np.uint64)
        if trial == 0:
            seeds[:] = 0
        if trial == 1:
            seeds[:] = 2**32 - 1
        gids = np.concatenate([special, rng.integers(0, 2**32, size=250, dtype=np.uint64)])
        want = K.hybrid(seeds, gids)
        got = np.array([O.hybrid_taus(seeds.astype(np.uint32), int(g)) for g in gids], dtype=np.float32)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), f"seeds {seeds}"
        n_checked += len(gids)
    assert n_checked >= 10_000
    # a hand-checked value: all-zero seeds at gid 0 leave only the LCG constant
    # and the three Tausworthe steps of 128
    w0 = K.hybrid_bits([0, 0, 0, 0], 0)[0]
    t = [K.taus_step(np.uint32(128), *c[1:])[()] for c in CONST["taus"]]
    assert w0 == np.uint32(CONST["lcg"][1]) ^ t[0] ^ t[1] ^ t[2]


def test_per_iteration_seed_sets_are_glibc_rand_in_order():
    """Iteration t of a run seeded with s reads rand() values 16(t+1) ..
    16(t+2)-1 as seeds[0].z1 .. seeds[3].z4 (src/abms.c:140-158, 511, 671):
    the oracle's stream against the host C library's."""
    from tests import kat_rng as K
    for seed in (1, 7, 123456):
        sets = K.seed_sets(seed, 5)
        flat = O.seed_stream(seed, 16 * 6)
        assert np.array_equal(sets.reshape(-1), flat[16:])
