"""The one-workgroup FP64 Rayleigh-Ritz eigensolver against numpy.linalg.eigh
(a textbook routine), on spectra shaped like the block eigensolver's:
clusters, exact multiplicities, and a spread of eleven orders of magnitude."""

import numpy as np
import pytest

from lapy_amd import _lib
from oracle import metrics

pytestmark = pytest.mark.gpu


def sym_with_spectrum(lam, seed):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((len(lam), len(lam))))
    return (q * lam) @ q.T


def check(H, nev, tol_val=1e-12):
    ctx = _lib.default_context()
    ev, Z, ph = _lib.small_symeig(ctx, H, nev, profile=True)
    w, V = np.linalg.eigh(H)
    nrm = np.abs(w).max()
    assert np.abs(ev - w[:nev]).max() <= tol_val * nrm, (ev - w[:nev])
    assert np.abs(Z.T @ Z - np.eye(nev)).max() < 1e-12
    R = H @ Z - Z * ev
    assert np.abs(R).max() <= 1e-11 * nrm
    # subspace check per cluster of the reference spectrum
    for g in metrics.clusters(w[:nev], cluster_rtol=1e-6):
        if g[-1] == nev - 1 and nev < len(w) and abs(w[nev] - w[nev - 1]) <= 1e-6 * max(abs(w[nev]), 1e-300):
            continue
        s = np.linalg.svd(V[:, g].T @ Z[:, g], compute_uv=False)
        assert s.min() > 1 - 1e-10
    return ph


# every register-layout class of the tridiagonalisation (ceil(N / 32) = 1 .. 6),
# at and off the class boundaries
@pytest.mark.parametrize("N,nev", [(192, 64), (168, 56), (161, 50), (160, 53), (150, 50), (128, 64), (96, 32),
                                   (65, 20), (64, 64), (48, 16), (33, 11), (17, 5), (3, 2), (2, 1), (1, 1)])
def test_random_spectrum(N, nev):
    lam = np.sort(np.random.default_rng(N).uniform(-3, 10, N))
    check(sym_with_spectrum(lam, N), nev)


def test_clusters_and_multiplicities():
    lam = np.concatenate([[0.0, 1e-13], np.repeat([2.0, 6.0], 5), 6.0 + 1e-9 * np.arange(7),
                          np.linspace(7, 1e3, 192 - 19)])
    check(sym_with_spectrum(np.sort(lam), 1), 64)


def test_diagonal_and_block_diagonal():
    """A diagonal matrix has a zero column below every diagonal entry (every
    reflector is the identity, tau = 0); two decoupled blocks give zeros in
    part of every column."""
    rng = np.random.default_rng(3)
    d = rng.uniform(-2.0, 5.0, 96)
    check(np.diag(d), 20)
    H = np.zeros((100, 100))
    H[:40, :40] = sym_with_spectrum(np.linspace(1.0, 2.0, 40), 5)
    H[40:, 40:] = sym_with_spectrum(np.linspace(0.5, 3.0, 60), 6)
    check(H, 30)


@pytest.mark.parametrize("scale", [1e-150, 1e150])
def test_extreme_magnitudes(scale):
    """Norms beyond the range of the short-latency square root and reciprocal
    take the exact ones; the bisection works on T scaled by a power of two."""
    lam = np.sort(np.random.default_rng(7).uniform(-3.0, 10.0, 168))
    check(sym_with_spectrum(lam, 9) * scale, 56)


def test_lobpcg_like_wide_spectrum():
    rng = np.random.default_rng(5)
    lam = np.sort(np.concatenate([[1e-12], rng.uniform(0.4, 60.0, 63), 10 ** rng.uniform(2, 5.7, 128)]))
    ph = check(sym_with_spectrum(lam, 2), 64, tol_val=2e-13)
    print("phase microseconds:", np.round(ph, 1), "total", round(float(ph.sum()), 1))


# ---- the Cholesky-QR factor (lh_small_cholinv) against numpy's Cholesky ------
def _spd(m, seed, spread=4.0):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((m + 10, m))
    sc = np.exp(rng.uniform(-spread, spread, m))
    return sc[:, None] * (a.T @ a) * sc[None, :]


@pytest.mark.parametrize("m", [1, 2, 5, 31, 32, 33, 63, 64, 65, 96, 100, 112, 127, 128])
def test_cholinv_is_the_inverse_cholesky_factor(m):
    """T = D L^-T for the Cholesky factor L of the equilibrated D G D (D =
    diag(G)^-1/2): upper triangular, T^T G T = I, and the smallest pivot is
    min diag(L)^2.  Odd m exercise the identity padding of the register
    layout, every ceil(m / 32) class its own instantiation."""
    ctx = _lib.default_context()
    G = _spd(m, m)
    T, mp = _lib.small_cholinv(ctx, G)
    assert np.all(np.tril(T, -1) == 0.0)
    d = 1.0 / np.sqrt(np.diag(G))
    L = np.linalg.cholesky(d[:, None] * G * d[None, :])
    Tref = d[:, None] * np.linalg.inv(L).T
    err = np.abs(T - Tref).max() / np.abs(Tref).max()
    print(f"m={m}: |T - D L^-T| / |T| = {err:.2e}, min pivot {mp:.6e}")
    assert err <= 1e-11
    np.testing.assert_allclose(T.T @ G @ T, np.eye(m), atol=1e-11)
    assert abs(mp - np.min(np.diag(L)) ** 2) <= 1e-12 * max(1.0, mp)


def test_cholinv_shift_and_rank_deficiency():
    """A shift s adds s to the equilibrated diagonal: T^T (G + s D^-2) T = I.
    A rank-deficient G is reported as LH_ERR_NUMERIC, and the shift makes it
    factorable."""
    ctx = _lib.default_context()
    G = _spd(40, 7)
    T, _ = _lib.small_cholinv(ctx, G, shift=0.25)
    dinv2 = np.diag(G)
    np.testing.assert_allclose(T.T @ (G + 0.25 * np.diag(dinv2)) @ T, np.eye(40), atol=1e-11)
    rng = np.random.default_rng(3)
    a = rng.standard_normal((30, 24))
    a[:, 23] = a[:, :23] @ rng.standard_normal(23)  # rank 23
    Gs = a.T @ a
    with pytest.raises(_lib.LapyHipError, match="positive definite"):
        _lib.small_cholinv(ctx, Gs)
    T, mp = _lib.small_cholinv(ctx, Gs, shift=1e-3)
    assert mp >= 1e-3 * (1 - 1e-12)
    np.testing.assert_allclose(T.T @ (Gs + 1e-3 * np.diag(np.diag(Gs))) @ T, np.eye(24), atol=1e-9)
