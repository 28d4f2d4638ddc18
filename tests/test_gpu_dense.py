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


@pytest.mark.parametrize("N,nev", [(192, 64), (128, 64), (96, 32), (17, 5), (3, 2), (1, 1)])
def test_random_spectrum(N, nev):
    lam = np.sort(np.random.default_rng(N).uniform(-3, 10, N))
    check(sym_with_spectrum(lam, N), nev)


def test_clusters_and_multiplicities():
    lam = np.concatenate([[0.0, 1e-13], np.repeat([2.0, 6.0], 5), 6.0 + 1e-9 * np.arange(7),
                          np.linspace(7, 1e3, 192 - 19)])
    check(sym_with_spectrum(np.sort(lam), 1), 64)


def test_lobpcg_like_wide_spectrum():
    rng = np.random.default_rng(5)
    lam = np.sort(np.concatenate([[1e-12], rng.uniform(0.4, 60.0, 63), 10 ** rng.uniform(2, 5.7, 128)]))
    ph = check(sym_with_spectrum(lam, 2), 64, tol_val=2e-13)
    print("phase microseconds:", np.round(ph, 1), "total", round(float(ph.sum()), 1))
