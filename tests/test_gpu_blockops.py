"""The eigensolver's tall-skinny FP64 block products (matrix-core kernels)
against numpy: Gram C = U^T V (full and symmetric-tile forms) and the update
Y = alpha U C + beta Y, over widths that are and are not multiples of the
64-wide tiles, narrow blocks (at most 9 useful 16 x 16 matrix-core blocks, the
Gram's narrow kernel), leading dimensions larger than the width, odd leading
dimensions (unaligned rows: the element-wise load path) and row counts that
are not multiples of the 16-row groups."""

import numpy as np
import pytest

from lapy_amd import _lib

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,p,q,ldu,ldv", [
    (10000, 64, 64, 192, 192),
    (10007, 64, 128, 192, 192),
    (4099, 192, 192, 192, 192),
    (3001, 56, 112, 168, 168),
    (2000, 100, 37, 101, 37),
    (17, 5, 3, 7, 3),
    # narrow blocks (at most 9 of the 16 matrix-core blocks of the tile used)
    (10242, 16, 16, 16, 16),
    (10242, 48, 48, 48, 48),
    (5001, 33, 47, 48, 47),
    (8000, 16, 48, 64, 48),
    (4000, 64, 32, 64, 32),
    (1001, 40, 20, 41, 21),
])
def test_gram_matches_numpy(n, p, q, ldu, ldv):
    rng = np.random.default_rng(n + p + q)
    U = rng.standard_normal((n, ldu))
    V = rng.standard_normal((n, ldv))
    C, _ = _lib.block_gram(_lib.default_context(), U, V, p, q)
    ref = U[:, :p].T @ V[:, :q]
    scale = np.sqrt(n)
    np.testing.assert_allclose(C, ref, rtol=0, atol=1e-13 * scale * 8)


@pytest.mark.parametrize("n,p,ld", [(10000, 64, 192), (5003, 128, 192), (4099, 192, 192), (1001, 72, 73),
                                    (10242, 48, 48), (999, 33, 35)])
def test_gram_symmetric_tiles(n, p, ld):
    """S^T (A S) with A SPD: the symmetric form fills the lower tiles from the
    upper ones, so tiles off the diagonal are exactly mirrored."""
    rng = np.random.default_rng(p)
    S = rng.standard_normal((n, ld))
    d = rng.uniform(0.5, 2.0, n)
    AS = S * d[:, None]
    C, _ = _lib.block_gram(_lib.default_context(), S, AS, p, p, sym=True)
    ref = S[:, :p].T @ AS[:, :p]
    np.testing.assert_allclose(C, ref, rtol=0, atol=1e-12 * np.sqrt(n) * 8)
    for i in range(p):
        for j in range(p):
            if i // 64 != j // 64:
                assert C[i, j] == C[j, i]


def test_gram_deterministic():
    rng = np.random.default_rng(3)
    U = rng.standard_normal((50000, 192))
    ctx = _lib.default_context()
    c1, _ = _lib.block_gram(ctx, U, U, 192, 192, sym=True)
    c2, _ = _lib.block_gram(ctx, U, U, 192, 192, sym=True)
    assert np.array_equal(c1, c2)


@pytest.mark.parametrize("n,p,q,ldu,ldy,alpha,beta", [
    (10000, 192, 64, 192, 192, 1.0, 0.0),
    (10007, 128, 128, 192, 128, 1.0, 0.0),
    (4099, 64, 128, 192, 192, -1.0, 1.0),
    (3001, 56, 112, 168, 168, 0.5, -2.0),
    (2000, 37, 100, 37, 101, 1.0, 0.0),
    (5003, 64, 192, 64, 192, 1.0, 1.0),
    (3001, 64, 256, 64, 256, -1.0, 0.5),
    (2049, 48, 300, 48, 300, 1.0, 0.0),
    (17, 3, 5, 3, 5, 2.0, 1.0),
    (3001, 56, 48, 56, 48, 1.0, 0.5),          # three 16-column tiles: 960-thread workgroups
    (40000, 192, 64, 192, 192, 1.0, 0.0),      # n large enough for 64-column tiles
    (33000, 128, 128, 192, 128, -1.0, 1.0),    # 64-column tiles, two per workgroup
])
def test_gemm_matches_numpy(n, p, q, ldu, ldy, alpha, beta):
    rng = np.random.default_rng(n * 7 + p)
    U = rng.standard_normal((n, ldu))
    Cm = rng.standard_normal((p, q))
    Y0 = rng.standard_normal((n, ldy))
    Y, _ = _lib.block_gemm(_lib.default_context(), U, p, Cm, Y0, alpha, beta)
    ref = Y0.copy()
    ref[:, :q] = alpha * (U[:, :p] @ Cm) + beta * Y0[:, :q]
    np.testing.assert_allclose(Y, ref, rtol=0, atol=1e-13 * np.sqrt(p) * 8 * (1 + abs(beta)))
    # columns beyond q are untouched
    assert np.array_equal(Y[:, q:], Y0[:, q:])


# small n run on 16-column tiles, n >= 32,768 (at 64 columns) on 64-column ones
@pytest.mark.parametrize("n,p,ld,upper", [(10000, 64, 192, True), (10007, 128, 192, True), (4099, 64, 192, False),
                                          (3001, 192, 192, False), (1001, 72, 73, True), (17, 5, 5, False),
                                          (40000, 128, 192, True), (40000, 64, 64, False)])
def test_gemm_in_place(n, p, ld, upper):
    """X <- X C in place (the Cholesky-QR and Ritz rotations): every
    workgroup reads all columns of its rows before any of them is written."""
    rng = np.random.default_rng(n + p)
    X0 = rng.standard_normal((n, ld))
    Cm = rng.standard_normal((p, p))
    if upper:
        Cm = np.triu(Cm)
    X, _ = _lib.block_gemm(_lib.default_context(), None, p, Cm, X0, 1.0, 0.0, upper=upper, inplace=True)
    ref = X0.copy()
    ref[:, :p] = X0[:, :p] @ Cm
    np.testing.assert_allclose(X, ref, rtol=0, atol=1e-13 * np.sqrt(p) * 8)
    assert np.array_equal(X[:, p:], X0[:, p:])


def test_gemm_upper_skips_only_zero_blocks():
    rng = np.random.default_rng(11)
    U = rng.standard_normal((5000, 128))
    Cm = np.triu(rng.standard_normal((128, 128)))
    Y0 = np.zeros((5000, 128))
    ctx = _lib.default_context()
    y_full, _ = _lib.block_gemm(ctx, U, 128, Cm, Y0)
    y_up, _ = _lib.block_gemm(ctx, U, 128, Cm, Y0, upper=True)
    np.testing.assert_allclose(y_up, y_full, rtol=0, atol=1e-12)


def _f32(a):
    return a.astype(np.float32).astype(np.float64)


@pytest.mark.parametrize("n,p,q,ld,sym", [
    (10000, 64, 64, 192, False),
    (10007, 192, 192, 192, True),
    (3001, 56, 112, 168, False),
    (2000, 100, 37, 101, False),
    (17, 5, 3, 7, False),
    (10242, 48, 48, 48, True),
    (3001, 16, 40, 41, False),
])
def test_gram_fp32_matrix_cores(n, p, q, ld, sym):
    """FP32 blocks (the mixed-precision eigensolver's): products of the
    FP32-rounded inputs, FP32 accumulation within a wave's rows (relative
    error ~1e-6), FP64 across waves."""
    rng = np.random.default_rng(n + 3 * p + q)
    U = rng.standard_normal((n, ld))
    V = U * rng.uniform(0.5, 2.0, n)[:, None] if sym else rng.standard_normal((n, ld))
    if sym:
        q = p
    C, _ = _lib.block_gram(_lib.default_context(), U, V, p, q, sym=sym, f32=True)
    ref = _f32(U)[:, :p].T @ _f32(V)[:, :q]
    np.testing.assert_allclose(C, ref, rtol=0, atol=2e-6 * np.sqrt(n) * 8)
    # an FP64 Gram of the unrounded inputs differs by the rounding, far more
    # than the FP32 Gram's own error: the rounding really happened
    assert np.abs(C - U[:, :p].T @ V[:, :q]).max() > 1e-10


@pytest.mark.parametrize("n,p,q,ld,alpha,beta,inplace,upper", [
    (10000, 192, 64, 192, 1.0, 0.0, False, False),
    (4099, 64, 128, 192, -1.0, 1.0, False, False),
    (2000, 37, 100, 101, 1.0, 0.0, False, False),
    (10007, 128, 128, 192, 1.0, 0.0, True, True),
    (1001, 72, 72, 73, 1.0, 0.0, True, False),
])
def test_gemm_fp32_matrix_cores(n, p, q, ld, alpha, beta, inplace, upper):
    rng = np.random.default_rng(n * 5 + p)
    U = rng.standard_normal((n, ld))
    Cm = rng.standard_normal((p, q))
    if upper:
        Cm = np.triu(Cm)
    Y0 = rng.standard_normal((n, ld))
    Y, _ = _lib.block_gemm(_lib.default_context(), None if inplace else U, p, Cm, Y0, alpha, beta, f32=True,
                           inplace=inplace, upper=upper)
    src = _f32(Y0) if inplace else _f32(U)
    ref = _f32(Y0)
    ref[:, :q] = alpha * (src[:, :p] @ _f32(Cm)) + beta * _f32(Y0)[:, :q]
    np.testing.assert_allclose(Y, ref, rtol=0, atol=3e-6 * np.sqrt(p) * 8 * (1 + abs(beta)))
    assert np.array_equal(Y[:, q:], _f32(Y0)[:, q:])
