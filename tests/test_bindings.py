"""The Python wrappers call every C entry point with arguments that ctypes
accepts under the declared signatures.  The library is replaced by ctypes
callbacks built from the same signature table, so the argument conversion is
exactly the one a real call performs (no GPU needed)."""

import ctypes as C
import types

import numpy as np
import pytest

from lapy_amd import _lib


@pytest.fixture
def fake(monkeypatch):
    calls = []
    ns = types.SimpleNamespace()
    keep = []
    for name, (res, args) in _lib.SIGNATURES.items():
        proto = C.CFUNCTYPE(res, *args)

        def cb(*a, _name=name, _res=res):
            calls.append(_name)
            return 0 if _res is not None else None

        f = proto(cb)
        keep.append(f)
        setattr(ns, name, f)
    monkeypatch.setattr(_lib, "lib", lambda: ns)
    ctx = types.SimpleNamespace(handle=None)
    yield ctx, calls


def test_block_products_argument_types(fake):
    ctx, calls = fake
    U = np.zeros((10, 6))
    _lib.block_gram(ctx, U, U, 4, 4, sym=True)
    _lib.block_gram(ctx, U, U, 4, 2, iters=3)
    _lib.block_gemm(ctx, U, 4, np.zeros((4, 3)), np.zeros((10, 5)), 1.0, 0.5)
    _lib.block_gemm(ctx, U, 4, np.zeros((4, 3)), np.zeros((10, 5)), iters=2)
    _lib.block_gemm(ctx, None, 4, np.zeros((4, 4)), np.zeros((10, 6)), inplace=True, upper=True)
    _lib.block_gram(ctx, U, U, 4, 4, f32=True)
    _lib.block_gemm(ctx, U, 4, np.zeros((4, 3)), np.zeros((10, 5)), f32=True)
    assert calls == ["lh_block_gram", "lh_block_gram", "lh_block_gemm", "lh_block_gemm", "lh_block_gemm",
                     "lh_block_gram", "lh_block_gemm"]


def test_eigs_and_timing_argument_types(fake):
    ctx, calls = fake
    A = types.SimpleNamespace(handle=None, n=10)
    _lib.eigs(ctx, A, A, 3, poly_storage=_lib.LH_STORE_F16, poly_degree=8)
    _lib.eigs(ctx, A, A, 3, keep_device=True)
    _lib.time_cheb_step(ctx, A, A, 64, iters=5, reorder=False, storage=_lib.LH_STORE_F32)
    assert calls.count("lh_eigs") == 2 and "lh_time_cheb_step" in calls


def test_batched_argument_types(fake):
    ctx, calls = fake
    A = types.SimpleNamespace(handle=None, n=10)
    res = _lib.eigs_batched(ctx, [A, A], [A, A], [3, 4], workers=2, poly_degree=8)
    assert res == [None, None]
    best, secs = _lib.eigs_autotune(ctx, [A], [A], [3])
    assert set(secs) == {1, 2, 4, 8, 16}
    assert calls == ["lh_eigs_batched", "lh_eigs_autotune"]


def test_mpnn_argument_types(fake):
    ctx, calls = fake
    A = types.SimpleNamespace(handle=None, n=6)
    net = _lib.Mpnn.__new__(_lib.Mpnn)
    net.ctx, net.handle = ctx, None
    net.apply(A, None, 0.01, 1.0, np.zeros((6, 2)))
    net.loss_grad(A, A, 0.01, 1.0, np.zeros((6, 2)))
    net.use()
    called = [c for c in calls if c != "lh_mpnn_param_count"]
    assert called[-3:] == ["lh_mpnn_apply", "lh_mpnn_loss_grad", "lh_ctx_set_mpnn"]
