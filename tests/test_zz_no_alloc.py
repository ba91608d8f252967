"""Allocation-free min cut (iq_graphcut_noalloc) against iq_graphcut.

The caller owns every device buffer (hipMalloc through ctypes here); the library keeps
the cut's status word and the node state of large slabs in its static
per-device array.  Two calls in a row, the second on a smaller slab, must
both equal iq_graphcut's masks bit for bit (and the oracle's), so nothing
the first call left in the array leaks into the second.  Slabs above the
array's limit are refused before any device work.
"""
import ctypes

import numpy as np
import pytest

import iqhip
import iqref
from iqhip import _lib

pytestmark = pytest.mark.gpu


def _smooth(rng, shape, width, dtype):
    x = rng.standard_normal(shape)
    k = np.ones(width) / width
    for ax in range(len(shape)):
        if shape[ax] > width:
            x = np.apply_along_axis(lambda v: np.convolve(v, k, mode="same"), ax, x)
    return x.astype(dtype)


def _dims(shape):
    s = tuple(shape) + (1,) * (3 - len(shape))
    return (ctypes.c_int64 * 3)(*s)


def _hip():
    """The HIP runtime libiqhip.so links (the caller's own allocator here)."""
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return h


def _ok(rc):
    assert rc == 0, f"HIP error {rc}"


def _cut_noalloc(ctx, A, B, dim):
    hip = _hip()
    dt = _lib.IQ_F32 if A.dtype == np.float32 else _lib.IQ_F64
    # column-major slabs: the Fortran-order bytes of A and B, in caller-owned device memory
    a = np.ascontiguousarray(A.ravel(order="F"))
    b = np.ascontiguousarray(B.ravel(order="F"))
    m = np.full(A.size, 7, dtype=np.uint8)
    dA, dB, dM = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    for p, n in ((dA, a.nbytes), (dB, b.nbytes), (dM, m.nbytes)):
        _ok(hip.hipMalloc(ctypes.byref(p), n))
    try:
        _ok(hip.hipMemcpy(dA, a.ctypes.data, a.nbytes, 1))  # host to device
        _ok(hip.hipMemcpy(dB, b.ctypes.data, b.nbytes, 1))
        _ok(hip.hipMemcpy(dM, m.ctypes.data, m.nbytes, 1))
        rc = ctx._L.iq_graphcut_noalloc(ctx.handle, dA, dB, dt, _dims(A.shape), int(dim), dM)
        ctx.check(rc)
        _ok(hip.hipMemcpy(m.ctypes.data, dM, m.nbytes, 2))  # device to host
    finally:
        for p in (dA, dB, dM):
            hip.hipFree(p)
    m = m.reshape(A.shape, order="F")
    assert set(np.unique(m).tolist()) <= {0, 1}
    return m.astype(bool)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_noalloc_cut_matches_graphcut_twice(ctx, dtype):
    rng = np.random.default_rng(41)
    # interior nodes: 10*32*8 = 2,560 (node state in the static array), then
    # 8*28*8 = 1,792 (smaller, same array), then 4*24*6 = 576 (LDS)
    cases = [((12, 32, 8), 0), ((10, 28, 8), 0), ((6, 24, 6), 0)]
    for shape, dim in cases:
        print("noalloc cut", dtype.__name__, shape, dim, flush=True)
        A = _smooth(rng, shape, 3, dtype)
        B = _smooth(rng, shape, 3, dtype)
        got = _cut_noalloc(ctx, A, B, dim)
        ref = iqhip.graphcut(A, B, dim, ctx=ctx)
        assert np.array_equal(got, ref), shape
        assert np.array_equal(got, iqref.graphcut(A, B, dim)), shape


def test_noalloc_cut_refuses_slabs_beyond_the_array(ctx):
    n = _lib.IQ_NOALLOC_MAX_NODES
    shape = (n // 64 + 3, 64, 1)  # interior (n/64 + 1) * 64 > n nodes
    dims = _dims(shape)
    dummy = ctypes.c_void_p(1)  # never dereferenced: the size check comes first
    rc = ctx._L.iq_graphcut_noalloc(ctx.handle, dummy, dummy, _lib.IQ_F32, dims, 0, dummy)
    assert rc == -6  # IQ_ERR_UNSUPPORTED
    assert b"at most" in ctx._L.iq_last_error(ctx.handle)
