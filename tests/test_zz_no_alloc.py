"""lh_block_gram_noalloc (the Gram product of the eigensolver without any
device allocation, its partials in a fixed per-GPU scratch array) against
lh_block_gram, through the C ABI, on two calls in a row -- the second with a
smaller input, so that stale scratch contents from the first call would
show -- and its size limit.

The device buffers of the inputs and outputs are made by this test with the
HIP runtime's own hipMalloc (the caller's memory); the entry point allocates
none.  The two entry points split the rows over different numbers of waves,
so they agree to rounding: |C_noalloc - C| <= 1e-13 * (|U|^T |V|) elementwise.
"""

import ctypes as C

import numpy as np
import pytest

from lapy_amd import _lib

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


def _hip():
    _lib.lib()  # the library first: the HIP runtime it links is the one loaded below
    h = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMalloc.restype = C.c_int
    h.hipFree.argtypes = [C.c_void_p]
    h.hipFree.restype = C.c_int
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemcpy.restype = C.c_int
    return h


class _Dev:
    """A device copy of a host array (hipMalloc + hipMemcpy), freed on close."""

    def __init__(self, hip, a):
        self.hip = hip
        self.nbytes = a.nbytes
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(self.nbytes, 8)) == 0
        assert hip.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), self.nbytes, _H2D) == 0

    def download(self, out):
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, out.nbytes, _D2H) == 0
        return out

    def close(self):
        self.hip.hipFree(self.p)


def _gram_noalloc(ctx, hip, U, V, p, q, sym):
    dU, dV = _Dev(hip, U), _Dev(hip, V)
    out = np.full((p, q), np.nan)
    dC = _Dev(hip, out)
    try:
        code = _lib.lib().lh_block_gram_noalloc(ctx.handle, U.shape[0], dU.p, U.shape[1], p, dV.p, V.shape[1], q,
                                                _lib.LH_GRAM_SYM if sym else 0, dC.p)
        _lib.check(code, ctx, "lh_block_gram_noalloc")
        ctx.sync()
        return dC.download(out)
    finally:
        for d in (dU, dV, dC):
            d.close()


def _check(ctx, hip, n, ldu, p, ldv, q, sym, seed):
    rng = np.random.default_rng(seed)
    U = rng.standard_normal((n, ldu))
    V = U if sym else rng.standard_normal((n, ldv))
    print(f"case: n={n} p={p} (ld {ldu}) q={q} (ld {V.shape[1]}) sym={sym}", flush=True)
    ref, _ = _lib.block_gram(ctx, U, V, p, q, sym=sym)
    got = _gram_noalloc(ctx, hip, U, V, p, q, sym)
    bound = np.abs(U[:, :p]).T @ np.abs(V[:, :q])
    diff = np.abs(got - ref)
    print(f"  max |diff| / bound {np.max(diff / bound):.2e}")
    assert np.all(np.isfinite(got))
    assert np.all(diff <= 1e-13 * bound)
    if sym:
        assert np.array_equal(got, got.T)


def test_noalloc_gram_matches_block_gram_twice_then_smaller():
    ctx = _lib.default_context()
    hip = _hip()
    # first call: the eigensolver's 192-column symmetric Gram (6 upper tiles)
    _check(ctx, hip, 200_003, 192, 192, 192, 192, True, 1)
    # second call, smaller and of another shape: ragged widths, padded rows
    _check(ctx, hip, 5_001, 72, 70, 40, 37, False, 2)


def test_noalloc_gram_rejects_sizes_beyond_its_scratch():
    ctx = _lib.default_context()
    hip = _hip()
    n = 64
    big = _lib.LH_GRAM_NOALLOC_MAX_COLS + 1
    U = np.zeros((n, big))
    dU = _Dev(hip, U)
    dC = _Dev(hip, np.zeros(8))
    try:
        code = _lib.lib().lh_block_gram_noalloc(ctx.handle, n, dU.p, big, big, dU.p, big, big, 0, dC.p)
        assert code == _lib.LH_ERR_ARG
        assert "scratch" in _lib.lib().lh_last_error(ctx.handle).decode()
        code = _lib.lib().lh_block_gram_noalloc(ctx.handle, n, dU.p, big, 8, dU.p, big, 8, _lib.LH_BLOCK_F32, dC.p)
        assert code == _lib.LH_ERR_ARG
    finally:
        dU.close()
        dC.close()
