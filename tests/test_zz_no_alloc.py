"""bhx_sort_keys_noalloc (the key sort on device arrays, with no allocation)
against bhx_sort_keys (host arrays, the context's own buffers): two calls in a
row on the same device, the second with a smaller input, then the argument
checks.  The caller's device arrays are allocated here with hipMalloc from the
HIP runtime the library itself is linked against; the library allocates
nothing for the call."""
import ctypes

import numpy as np
import pytest

from tests.helpers import load_bhx

pytestmark = pytest.mark.gpu

MAX_KEYS = 1 << 20
H2D, D2H = 1, 2     # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _hip():
    """The HIP runtime libbhx.so uses (loaded with it, found by its soname)."""
    load_bhx().lib()
    H = ctypes.CDLL("libamdhip64.so.7")
    P, S = ctypes.c_void_p, ctypes.c_size_t
    H.hipMalloc.argtypes = [ctypes.POINTER(P), S]
    H.hipFree.argtypes = [P]
    H.hipMemcpy.argtypes = [P, P, S, ctypes.c_int]
    H.hipGetDevice.argtypes = [ctypes.POINTER(ctypes.c_int)]
    for f in (H.hipMalloc, H.hipFree, H.hipMemcpy, H.hipGetDevice):
        f.restype = ctypes.c_int
    return H


class _DevArray:
    """A device copy of a uint64 host array (caller-owned memory)."""

    def __init__(self, H, host):
        self.H = H
        self.n = host.size
        self.ptr = ctypes.c_void_p()
        assert H.hipMalloc(ctypes.byref(self.ptr), max(8, 8 * self.n)) == 0
        a = np.ascontiguousarray(host, dtype=np.uint64)
        assert H.hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), 8 * self.n, H2D) == 0

    def get(self):
        out = np.empty(self.n, np.uint64)
        assert self.H.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, 8 * self.n, D2H) == 0
        return out

    def free(self):
        if self.ptr:
            self.H.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def _keys(n, seed):
    rng = np.random.default_rng(seed)
    base = rng.integers(0, 2 ** 63, 257, dtype=np.int64).astype(np.uint64) | np.uint64(1 << 63)
    keys = base[rng.integers(0, 257, n)]                    # many equal keys: stability shows
    idx = rng.integers(0, 2 ** 62, n, dtype=np.int64).astype(np.uint64)
    return keys, idx


def test_sort_keys_noalloc_matches_sort_keys():
    bhx = load_bhx()
    L = bhx.lib()
    H = _hip()
    dev = ctypes.c_int()
    assert H.hipGetDevice(ctypes.byref(dev)) == 0
    with bhx.Engine(MAX_KEYS) as e:
        for n, seed in ((MAX_KEYS, 1), (100003, 2)):
            keys, idx = _keys(n, seed)
            ref_k, ref_i = e.sort_keys(keys, idx)
            dk, di = _DevArray(H, keys), _DevArray(H, idx)
            try:
                print("noalloc sort n=%d" % n, flush=True)
                rc = L.bhx_sort_keys_noalloc(dev.value, None, dk.ptr, di.ptr, n)
                assert rc == 0
                got_k, got_i = dk.get(), di.get()
            finally:
                dk.free()
                di.free()
            assert np.array_equal(got_k, ref_k) and np.array_equal(got_i, ref_i)
            assert np.array_equal(got_k, keys[np.argsort(keys, kind="stable")])


def test_sort_keys_noalloc_argument_errors():
    bhx = load_bhx()
    L = bhx.lib()
    H = _hip()
    dev = ctypes.c_int()
    assert H.hipGetDevice(ctypes.byref(dev)) == 0
    d = _DevArray(H, np.zeros(4, np.uint64))
    try:
        assert L.bhx_sort_keys_noalloc(dev.value, None, None, d.ptr, 4) == 1                 # BHX_E_NULL
        assert L.bhx_sort_keys_noalloc(dev.value, None, d.ptr, d.ptr, MAX_KEYS + 1) == 9     # BHX_E_CAPACITY
        assert L.bhx_sort_keys_noalloc(-1, None, d.ptr, d.ptr, 4) == 2                       # BHX_E_ARG
        assert L.bhx_sort_keys_noalloc(dev.value, None, d.ptr, d.ptr, 1) == 0                # nothing to sort
    finally:
        d.free()
