"""The compaction index (include/sampo_compact.h): every agent's rank among
the agents of its new state, which the reference gets from preScan and a
segmented scan per group of states (src/abms.c:359-415).

CPU: the numpy rank used as the expectation below is pinned to the oracle's
restatement of that reference scan (preScan flags + the three Bolt
scan-by-key kernels, inclusive minus one).  GPU: the HIP kernels equal it
bit for bit at edge sizes (tile and round boundaries), with more tiles than
the tile scan has threads, on unaligned device pointers, on a reused
workspace, and through size-independent properties at full size.
"""
import numpy as np
import pytest

import sampo_amd as S
from oracle import pyoracle as O


def expected(keys):
    """rank of each key among equal earlier keys (keys 0..7), NONE otherwise"""
    keys = np.asarray(keys, np.uint8)
    dest = np.full(keys.shape[0], S.COMPACT_NONE, np.uint32)
    counts = np.zeros(8, np.uint32)
    for k in range(8):
        idx = np.nonzero(keys == k)[0]
        dest[idx] = np.arange(idx.shape[0], dtype=np.uint32)
        counts[k] = idx.shape[0]
    return dest, counts


def random_keys(n, seed, dead=0.05):
    rng = np.random.default_rng(seed)
    k = rng.integers(0, 8, n).astype(np.uint8)
    k[rng.uniform(size=n) < dead] = 8
    return k


@pytest.mark.parametrize("n,seed", [(1, 1), (300, 2), (5000, 3), (20000, 4)])
def test_rank_is_the_reference_scan(n, seed):
    """preScan sets data = 1 for agents of the matched state (kernel/preScan.cl:57-61);
    the Bolt scan-by-key over one segment returns the inclusive sum minus one."""
    keys = random_keys(n, seed)
    dest, counts = expected(keys)
    for k in range(8):
        data = (keys == k).astype(np.int32)
        pre = O.bolt_scan(np.zeros(n, np.uint32), data, 128)
        idx = keys == k
        assert np.array_equal(pre[idx].astype(np.uint32), dest[idx]), k
        assert counts[k] == (pre[-1] + 1 if n else 0)


def test_workspace_size():
    # 32 bytes per tile of 16384 keys and per block of 1024 tiles
    assert S.compact_workspace_bytes(0) == 0
    assert S.compact_workspace_bytes(1) == 64
    assert S.compact_workspace_bytes(16384) == 64
    assert S.compact_workspace_bytes(16385) == 96
    assert S.compact_workspace_bytes(1024 * 16384 + 1) == 32 * (1025 + 2)


# ----------------------------------------------------------------- GPU


@pytest.mark.gpu
@pytest.mark.parametrize("n", [0, 1, 15, 16, 17, 4095, 4096, 4097, 16383, 16384, 16385, 32767,
                               256 * 16384 + 5, 600 * 16384 - 3, 1_000_003])
def test_gpu_compact_matches(n):
    keys = random_keys(n, n + 7)
    dest, counts = S.compact_index(keys)
    want_d, want_c = expected(keys)
    assert np.array_equal(counts, want_c)
    assert np.array_equal(dest, want_d)


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["all_dead", "one_state", "sorted", "alternating", "absent_keys"])
def test_gpu_compact_special_inputs(case):
    n = 16384 * 1100 + 11  # more than one block of 1024 tiles
    if case == "all_dead":
        keys = np.full(n, 8, np.uint8)
    elif case == "one_state":
        keys = np.full(n, 5, np.uint8)
    elif case == "sorted":
        keys = np.sort(random_keys(n, 3))
    elif case == "alternating":
        keys = (np.arange(n) % 8).astype(np.uint8)
    else:  # any key >= 8 is not ranked
        keys = random_keys(n, 4)
        keys[::7] = 200
    dest, counts = S.compact_index(keys)
    want_d, want_c = expected(keys)
    assert np.array_equal(counts, want_c) and np.array_equal(dest, want_d)


def hip_runtime():
    """The HIP runtime the library links (loaded with it), for device buffers."""
    import ctypes as C
    S.lib()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    return h


@pytest.mark.gpu
def test_gpu_compact_device_pointers_unaligned_and_workspace_reuse():
    import ctypes as C
    h = hip_runtime()
    H2D, D2H = 1, 2

    def alloc(nbytes):
        p = C.c_void_p()
        assert h.hipMalloc(C.byref(p), nbytes) == 0 and h.hipMemset(p, 0, nbytes) == 0
        return p.value

    sizes = [300_001, 4097, 123_457]  # one workspace, calls in a row, smaller after larger
    ws_bytes = S.compact_workspace_bytes(max(sizes))
    ws = alloc(ws_bytes)
    kbuf, dbuf, cbuf = alloc(max(sizes) + 16), alloc(4 * (max(sizes) + 8)), alloc(64)
    try:
        for i, n in enumerate(sizes):
            keys = random_keys(n, 40 + i)
            assert h.hipMemset(dbuf, 0, 4 * (n + 8)) == 0
            assert h.hipMemcpy(kbuf + 1, keys.ctypes.data, n, H2D) == 0   # keys at an odd address
            S.compact_index_device(kbuf + 1, n, dbuf + 4, cbuf, ws, ws_bytes)  # dest 4 bytes past alignment
            assert h.hipDeviceSynchronize() == 0
            out = np.zeros(n + 2, np.uint32)
            counts = np.zeros(8, np.uint32)
            assert h.hipMemcpy(out.ctypes.data, dbuf, 4 * (n + 2), D2H) == 0
            assert h.hipMemcpy(counts.ctypes.data, cbuf, 32, D2H) == 0
            want_d, want_c = expected(keys)
            assert np.array_equal(counts, want_c), n
            assert np.array_equal(out[1:n + 1], want_d), n
            assert out[0] == 0 and out[n + 1] == 0  # nothing written outside
    finally:
        for p in (ws, kbuf, dbuf, cbuf):
            h.hipFree(p)


@pytest.mark.gpu
def test_gpu_compact_full_size_properties():
    """n = 2^26: counts sum to the ranked keys and each key's ranks run
    0, 1, 2, ... in index order (stable)."""
    n = 1 << 26
    keys = random_keys(n, 99)
    dest, counts = S.compact_index(keys)
    assert int(counts.sum()) == int((keys < 8).sum())
    assert np.all(dest[keys >= 8] == S.COMPACT_NONE)
    for k in range(8):
        r = dest[keys == k]
        assert r.shape[0] == counts[k] and np.array_equal(r, np.arange(counts[k], dtype=np.uint32)), k
