"""Known-answer restatement of the reference's random streams, for pinning
the oracle (test helper, not a test module).

Everything here is built from numpy and from data, never from the oracle's
or the product's code:
  - HybridTaus (include/gpuRand.h:14-50): the Tausworthe shifts and masks,
    the LCG multiplier and increment, and how each component's state is
    formed from the seed and the global id are read from
    tests/golden/constants.json ("taus", "lcg", "hybridState"), which
    tests/golden/make_constants.py parses out of the reference header;
  - the per-iteration seed sets (src/abms.c:140-158): the host C library's
    own srand()/rand(), called through ctypes, 16 values per set in the order
    seeds[0].z1..z4, seeds[1].z1..z4, ..., seeds[3].z4.
"""
import ctypes
import json
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CONST = json.load(open(os.path.join(ROOT, "tests", "golden", "constants.json")))
TWO_M32 = np.float32(2.0 ** -32)


def taus_step(z, s1, s2, s3, m):
    """((z & M) << S3) ^ (((z << S1) ^ z) >> S2) on uint32 arrays."""
    z = np.asarray(z, dtype=np.uint32)
    b = np.right_shift(np.bitwise_xor(np.left_shift(z, np.uint32(s1)), z), np.uint32(s2))
    return np.bitwise_xor(np.left_shift(np.bitwise_and(z, np.uint32(m)), np.uint32(s3)), b)


def hybrid_bits(seed4, gid):
    """The 32-bit word HybridTaus scales by 2^-32, for seeds (z1..z4) and an
    array of global ids."""
    gid = np.atleast_1d(np.asarray(gid, dtype=np.uint64)).astype(np.uint32)
    z = {}
    for d in CONST["hybridState"]:
        v = np.full(gid.shape, np.uint32(int(seed4[d["z"] - 1]) & 0xFFFFFFFF), dtype=np.uint32)
        v = v + np.uint32(d["add"])
        if d["add_gid"]:
            v = v + gid
        if d["mul_gid"]:
            v = v * gid
        z[d["z"]] = v
    a, c = CONST["lcg"]
    out = np.uint32(a) * z[4] + np.uint32(c)
    for comp, s1, s2, s3, m in CONST["taus"]:
        out = np.bitwise_xor(out, taus_step(z[comp], s1, s2, s3, m))
    return out


def hybrid(seed4, gid):
    """HybridTaus as binary32: float(word) * 2^-32 (round to nearest, then an
    exact power-of-two scaling)."""
    return hybrid_bits(seed4, gid).astype(np.float32) * TWO_M32


_libc = None


def glibc_values(seed, n):
    """The first n values of the host C library's rand() after srand(seed)."""
    global _libc
    if _libc is None:
        _libc = ctypes.CDLL("libc.so.6")
    _libc.srand(ctypes.c_uint(seed))
    return np.array([_libc.rand() for _ in range(n)], dtype=np.uint32)


def seed_sets(seed, iterations):
    """sets[t][k] = seeds[k] (z1..z4) of iteration t of a run seeded with
    `seed`: set 0 is spent before the loop (src/abms.c:671), iteration t uses
    set t + 1 (src/abms.c:511)."""
    v = glibc_values(seed, 16 * (iterations + 1)).reshape(iterations + 1, 4, 4)
    return v[1:]
