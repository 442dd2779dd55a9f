"""The compaction index without allocation (sampo_compact_index_noalloc)
against the workspace entry point it varies (sampo_compact_index), through
the C ABI (MI355X).

Two calls in a row, the second on fewer keys, so that the second runs on a
fixed workspace that still holds the first call's contents; both must equal
the workspace entry point bit for bit (and the numpy rank).  Sizes beyond
the fixed workspace are refused before anything runs.  The test's own
buffers are allocated here with the HIP runtime; the entry point under test
allocates nothing.
"""
import ctypes as C

import numpy as np
import pytest

import sampo_amd as S

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Hip:
    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipSetDevice.argtypes = [C.c_int]
        self.bufs = []
        assert self.L.hipSetDevice(0) == 0

    def malloc(self, n):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(int(n), 1)) == 0
        self.bufs.append(p.value)
        return p.value

    def put(self, dst, arr):
        arr = np.ascontiguousarray(arr)
        assert self.L.hipMemcpy(dst, arr.ctypes.data, arr.nbytes, H2D) == 0

    def get(self, src, arr):
        assert self.L.hipDeviceSynchronize() == 0
        assert self.L.hipMemcpy(arr.ctypes.data, src, arr.nbytes, D2H) == 0
        return arr

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def lib():
    L = S.lib()
    L.sampo_compact_index.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                                      C.c_void_p]
    L.sampo_compact_index_noalloc.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p]
    return L


def expected(keys):
    dest = np.full(keys.shape[0], S.COMPACT_NONE, np.uint32)
    counts = np.zeros(8, np.uint32)
    for k in range(8):
        idx = np.nonzero(keys == k)[0]
        dest[idx] = np.arange(idx.shape[0], dtype=np.uint32)
        counts[k] = idx.shape[0]
    return dest, counts


def test_no_alloc_index_equals_the_workspace_entry_point():
    L = lib()
    hip = Hip()
    rng = np.random.default_rng(11)
    try:
        for n in (20_000_017, 1_234_567, 0):  # each call smaller than the last
            print(f"no-alloc compaction index on {n} keys", flush=True)
            keys = rng.integers(0, 9, n).astype(np.uint8)
            dk = hip.malloc(n)
            if n:
                hip.put(dk, keys)
            d1, d2 = hip.malloc(4 * n), hip.malloc(4 * n)
            c1, c2 = hip.malloc(32), hip.malloc(32)
            wsb = S.compact_workspace_bytes(n)
            ws = hip.malloc(wsb)
            assert L.sampo_compact_index(dk, n, d1, c1, ws, wsb, None) == 0
            assert L.sampo_compact_index_noalloc(dk, n, d2, c2, None) == 0
            got1 = hip.get(d1, np.zeros(n, np.uint32))
            got2 = hip.get(d2, np.zeros(n, np.uint32))
            cnt1 = hip.get(c1, np.zeros(8, np.uint32))
            cnt2 = hip.get(c2, np.zeros(8, np.uint32))
            assert np.array_equal(got1, got2) and np.array_equal(cnt1, cnt2)
            want, wc = expected(keys)
            assert np.array_equal(got2, want) and np.array_equal(cnt2, wc)
        # beyond the fixed workspace: refused on the host, nothing launched
        assert L.sampo_compact_index_noalloc(dk, (1 << 28) + 1, d2, c2, None) == -1  # SAMPO_E_ARG
    finally:
        hip.free()
