"""vh_fo_flag_noalloc (the allocation-free variant of vh_fo_flag, with the
flag kernels' per-block counts in the library's fixed device array) against
vh_fo_flag on the same two slices, through the C ABI: the same ordered cube
ids, bit for bit, on two calls in a row, the second with a smaller lattice
(so the array holds the first call's counts when the second starts).  The
caller's device buffers come from the HIP runtime the library itself uses
(hipMalloc through ctypes)."""
import ctypes

import numpy as np
import pytest

import cpuvh_amd as C
import vhtest as V
from cpuvh_amd import Hydro



def _slices(nx, ny, nz, seed):
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) - nx / 2) / nx
    y = (np.arange(ny) - ny / 2) / ny
    z = (np.arange(nz) - nz / 2) / nz if nz > 1 else np.zeros(1)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    base = 4.0 * np.exp(-(X * X + Y * Y + Z * Z) * 6.0)
    prev = (base * (1 + 0.05 * rng.standard_normal(base.shape))).ravel()
    cur = (0.8 * base * (1 + 0.05 * rng.standard_normal(base.shape))).ravel()
    return prev, cur


def _reference_ids(nx, ny, nz, prev, cur, eF):
    g = Hydro(nx, ny, nz, 0.1, 0.1, 0.1, 0.02, 0.6, 0.2)
    g.set("e", prev)
    g.fo_flag_ids(eF)  # previous := 0, current := prev
    g.set("e", cur)
    ids = g.fo_flag_ids(eF)  # previous := prev, current := cur
    g.close()
    return ids


class _Dev:
    """hipMalloc'd device buffers of the runtime libcpuvh_amd.so links."""

    def __init__(self):
        C._lib()
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.ptrs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), nbytes) == 0
        self.ptrs.append(p)
        return p

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def download(self, p, n, dtype):
        out = np.empty(n, dtype=dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host (synchronising)
        return out

    def close(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


@pytest.mark.gpu
def test_noalloc_flags_equal_vh_fo_flag_twice():
    L = C._lib()
    dev = _Dev()
    eF = 1.7
    for nx, ny, nz in [(61, 53, 17), (29, 23, 1)]:
        print(f"case noalloc fo flag {nx}x{ny}x{nz}", flush=True)
        prev, cur = _slices(nx, ny, nz, seed=nx)
        want = _reference_ids(nx, ny, nz, prev, cur, eF)
        assert want.size > 100
        M = (nx - 1) * (ny - 1) * max(nz - 1, 1)
        d_prev, d_cur = dev.upload(prev), dev.upload(cur)
        d_ids = dev.upload(np.full(M, -1, dtype=np.int64))
        d_n = dev.upload(np.full(1, -1, dtype=np.int64))
        st = L.vh_fo_flag_noalloc(d_prev, d_cur, nx, ny, nz, eF, d_ids, M, d_n, None)
        assert st == C.VH_OK
        n = int(dev.download(d_n, 1, np.int64)[0])
        ids = dev.download(d_ids, M, np.int64)
        assert n == want.size
        assert np.array_equal(ids[:n], want)
        assert (ids[n:] == -1).all()
        # a capacity below the count: the full count, only cap ids written
        cap = n // 2
        d_ids2 = dev.upload(np.full(M, -1, dtype=np.int64))
        st = L.vh_fo_flag_noalloc(d_prev, d_cur, nx, ny, nz, eF, d_ids2, cap, d_n, None)
        assert st == C.VH_OK
        ids = dev.download(d_ids2, M, np.int64)
        assert int(dev.download(d_n, 1, np.int64)[0]) == n
        assert np.array_equal(ids[:cap], want[:cap]) and (ids[cap:] == -1).all()
    dev.close()


def test_noalloc_rejects_oversize():
    """Beyond VH_FO_NOALLOC_MAX_CUBES cubes the call is refused before any
    device work (host only: the pointers are never used)."""
    L = C._lib()
    dummy = ctypes.c_void_p(16)
    st = L.vh_fo_flag_noalloc(dummy, dummy, 5001, 5001, 2, 1.0, dummy, 1, dummy, None)  # 25,000,000 cubes
    assert st == C.VH_ERR_ARG
    st = L.vh_fo_flag_noalloc(dummy, dummy, 1, 5, 5, 1.0, dummy, 1, dummy, None)  # nx < 2
    assert st == C.VH_ERR_ARG
