"""GPU: srcwa_zgeev_batched_noalloc, the eigensolve entry point that allocates nothing, against
srcwa_zgeev_batched through the C-ABI: bitwise the same eigenpairs and info on two calls in a row, the
second with a smaller batch (the shared scratch keeps the first call's contents), in place (V = A), and a
clear refusal of a batch larger than the scratch.

The test's own device buffers come from the HIP runtime the library links (libamdhip64.so.7, found by
soname once the library is loaded), never from torch, which ships a runtime of its own."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    from stablercwa_amd import _lib
    _lib.load()
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMalloc.restype = ctypes.c_int
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemcpy.restype = ctypes.c_int
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipFree.restype = ctypes.c_int
    return hip


class _Dev:
    """A device buffer of the library's HIP runtime."""

    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), max(int(nbytes), 16)) == 0
        self.nbytes = int(nbytes)

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = None


def _crand(rng, n):
    return rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))


def _noalloc(ctx, hip, mats, in_place=False):
    from stablercwa_amd import _lib
    from stablercwa_amd.rcwa import _pack
    n = np.array([m.shape[0] for m in mats], dtype=np.int32)
    A = _pack(mats)
    tot, nw = int(np.sum(n.astype(np.int64) ** 2)), int(n.sum())
    dA, dw, dinfo = _Dev(hip, A.nbytes), _Dev(hip, 16 * nw), _Dev(hip, 4 * len(mats))
    dV = dA if in_place else _Dev(hip, 16 * tot)
    try:
        dA.put(A)
        rc = _lib.load().srcwa_zgeev_batched_noalloc(ctx.handle, len(mats), _lib.ptr(n), dA.p, dw.p, dV.p, dinfo.p)
        _lib.check(rc)
        V = dV.get(np.complex128, tot)
        w = dw.get(np.complex128, nw)
        info = dinfo.get(np.int32, len(mats))
        A_after = None if in_place else dA.get(np.float64, A.size)
    finally:
        for b in {id(x): x for x in (dA, dw, dinfo, dV)}.values():
            b.free()
    out, a, b = [], 0, 0
    for i, k in enumerate(n):
        out.append((w[b:b + k], V[a:a + k * k].reshape((k, k), order="F"), int(info[i])))
        a += k * k
        b += k
    return out, A, A_after


def _same(ctx, hip, mats, in_place=False):
    from stablercwa_amd import rcwa
    ref = rcwa.zgeev_batched(mats, ctx)
    got, A, A_after = _noalloc(ctx, hip, mats, in_place)
    for (w1, V1, i1), (w2, V2, i2) in zip(ref, got):
        assert i1 == i2 == 0
        assert np.array_equal(w1, w2)
        assert np.array_equal(V1, V2)
    if A_after is not None:
        assert np.array_equal(A_after, A)          # the input is only read


def test_noalloc_equals_allocating_entry_point_twice(ctx):
    hip = _hip()
    rng = np.random.default_rng(31)
    _same(ctx, hip, [_crand(rng, n) for n in (121, 64, 200, 30, 7)])
    _same(ctx, hip, [_crand(rng, n) for n in (50, 17)])           # second call, smaller batch


def test_noalloc_in_place(ctx):
    hip = _hip()
    rng = np.random.default_rng(32)
    _same(ctx, hip, [_crand(rng, n) for n in (90, 41)], in_place=True)


def test_noalloc_refuses_batch_above_scratch(ctx):
    from stablercwa_amd import _lib
    hip = _hip()
    n = np.full(9, 1024, dtype=np.int32)          # 9 x 16 (2 n^2 + 2 n) bytes: above 64 MiB
    d = _Dev(hip, 16)
    try:
        rc = _lib.load().srcwa_zgeev_batched_noalloc(ctx.handle, len(n), _lib.ptr(n), d.p, d.p, d.p, d.p)
    finally:
        d.free()
    assert rc == _lib.ENOMEM
    assert b"split the batch" in _lib.load().srcwa_last_error()
