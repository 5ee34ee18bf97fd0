"""mf_knn_noalloc (the kNN search with its intermediates in a static device
array, for callers that cannot allocate) against mf_knn through the C ABI:
the same stencils and distances bit for bit on two calls in a row, the
second on a smaller cloud (the array keeps the first call's contents), and
a clear error for clouds beyond its 262,144-point limit."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from mfrbf import _lib as L

pytestmark = pytest.mark.gpu


def _hip():
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return h


def _knn_device(ctx, hip, N, k):
    d_idx, d_dist = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(d_idx), N * k * 4) == 0
    assert hip.hipMalloc(C.byref(d_dist), N * k * 8) == 0
    try:
        # poison the outputs so that unwritten entries show
        assert hip.hipMemset(d_idx, 0xff, N * k * 4) == 0
        assert hip.hipMemset(d_dist, 0xff, N * k * 8) == 0
        ctx.knn_noalloc(k, d_idx.value, d_dist.value)
        idx = np.empty((N, k), np.int32)
        dist = np.empty((N, k))
        assert hip.hipMemcpy(idx.ctypes.data, d_idx, N * k * 4, 2) == 0  # device to host
        assert hip.hipMemcpy(dist.ctypes.data, d_dist, N * k * 8, 2) == 0
    finally:
        hip.hipFree(d_idx)
        hip.hipFree(d_dist)
    return idx, dist


def test_knn_noalloc_matches_knn(mf):
    hip = _hip()
    for side, k, seed in ((200, 30, 20261202), (72, 15, 20261209)):  # the second call smaller
        print("knn_noalloc side=%d k=%d" % (side, k), flush=True)
        x, y, _, _, _ = O.cloud_square(side, seed)
        N = len(x)
        ctx = L.Context(0)
        ctx.set_points(x, y)
        idx, dist, _, _ = ctx.knn(k, want_dist=True)
        nidx, ndist = _knn_device(ctx, hip, N, k)
        assert np.array_equal(nidx, idx)
        assert np.array_equal(ndist, dist)


def test_knn_noalloc_rejects_large_clouds(mf):
    x, y, _, _, _ = O.cloud_square(513, 7)  # 263,169 points
    ctx = L.Context(0)
    ctx.set_points(x, y)
    with pytest.raises(L.MfError) as e:
        ctx.knn_noalloc(30, 1, 1)  # rejected before the addresses are used
    assert e.value.code == L.MF_ERR_UNSUPPORTED
    assert "262,144" in str(e.value)
