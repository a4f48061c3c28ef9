"""sc_kmeans_step_device_noalloc: the Lloyd step whose scratch is a fixed
device array (no hipMalloc) must give exactly what sc_kmeans_step gives, on
two calls in a row (the second on a smaller problem, so stale scratch from
the first call would show), and must refuse sizes beyond its scratch."""
import ctypes as C

import numpy as np
import pytest

from smartcore_amd import _capi as S
from smartcore_amd import synth

pytestmark = pytest.mark.gpu


class _Hip:
    """The HIP runtime the library itself loaded (by its soname), for the
    test's own device buffers: the test allocates and copies, the entry
    point under test only computes."""

    def __init__(self):
        S.lib()
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(nbytes, 16)) == 0
        return p

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0  # H2D
        return p

    def get(self, p, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0  # D2H
        return out


def _noalloc_step(ix, cent):
    hip = _Hip()
    L = S.lib()
    L.sc_kmeans_step_device_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p,
                                                C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    k, d = cent.shape
    n = ix.n
    # outputs prefilled with garbage: the call must overwrite all of them
    bufs = [hip.put(np.ascontiguousarray(cent, dtype=np.float64)),
            hip.put(np.full((k, d), 7.5)),
            hip.put(np.full(k, 12345, np.uint64)),
            hip.put(np.full(max(n, 1), 0x7777, np.uint32)),
            hip.put(np.full(1, -1.0))]
    try:
        rc = L.sc_kmeans_step_device_noalloc(ix.handle, bufs[0], k, bufs[1], bufs[2], bufs[3],
                                             bufs[4], None)
        S.check(rc)
        assert hip.L.hipDeviceSynchronize() == 0
        return (float(hip.get(bufs[4], 1, np.float64)[0]), hip.get(bufs[1], (k, d), np.float64),
                hip.get(bufs[2], k, np.uint64), hip.get(bufs[3], max(n, 1), np.uint32)[:n])
    finally:
        for b in bufs:
            hip.L.hipFree(b)


def _same_as_step(x, init):
    ix = S.Index(x)
    try:
        got = _noalloc_step(ix, init)
        want = ix.kmeans_step(init)
    finally:
        ix.close()
    assert got[0] == want[0] or (np.isnan(got[0]) and np.isnan(want[0]))
    np.testing.assert_array_equal(got[1].view(np.uint64), want[1].view(np.uint64))
    np.testing.assert_array_equal(got[2], want[2])
    np.testing.assert_array_equal(got[3], want[3])


def test_noalloc_step_equals_step_twice_second_smaller():
    x1, _, init1 = synth.kmeans_blobs(seed=31, n=50_000, d=24, blobs=12, k=12)
    _same_as_step(x1, init1)
    x2, _, init2 = synth.kmeans_blobs(seed=32, n=7_000, d=24, blobs=5, k=5)
    _same_as_step(x2, init2)
    # f64 rows and a dimension that is not a multiple of the row padding
    x3, _, init3 = synth.kmeans_blobs(seed=33, n=3_000, d=5, blobs=3, k=3, dtype=np.float64)
    _same_as_step(x3, init3)


def test_noalloc_step_refuses_sizes_beyond_scratch():
    x = np.random.default_rng(5).normal(size=(300_000, 64)).astype(np.float32)
    cent = x[:1024].astype(np.float64)
    ix = S.Index(x)
    try:
        with pytest.raises(S.Failed) as e:
            _noalloc_step(ix, cent)
        assert e.value.code == S.SC_ERR_UNSUPPORTED
        assert "scratch" in e.value.msg
    finally:
        ix.close()
