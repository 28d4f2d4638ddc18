"""Comparison measures between an eigen-result and the FP64 oracle.

Test infrastructure only (see ``oracle/__init__``).

* eigenvalue error: ``|lam - lam*| / |lam*|``; for a (near) zero eigenvalue
  the relative error is undefined, so ``|lam - lam*| / lam*_max`` is used for
  entries with ``|lam*| < 1e-6 * lam*_max``.
* eigenvectors: vectors are M-orthonormal with arbitrary sign and, inside a
  cluster of (nearly) equal eigenvalues, arbitrary basis.  They are compared
  per cluster by the largest principal angle between the two M-orthonormal
  bases; clusters are maximal runs of eigenvalues whose consecutive relative
  gap is below ``cluster_rtol``.
"""

from __future__ import annotations

import numpy as np


def eig_rel_err(lam, lam_ref):
    lam = np.asarray(lam, dtype=np.float64)
    lam_ref = np.asarray(lam_ref, dtype=np.float64)
    scale = np.max(np.abs(lam_ref))
    den = np.where(np.abs(lam_ref) < 1e-6 * scale, scale, np.abs(lam_ref))
    return np.abs(lam - lam_ref) / den


def clusters(lam_ref, cluster_rtol=1e-3):
    """Index groups of consecutive eigenvalues with relative gap < cluster_rtol."""
    lam_ref = np.asarray(lam_ref, dtype=np.float64)
    scale = max(np.max(np.abs(lam_ref)), 1e-300)
    groups = [[0]]
    for i in range(1, len(lam_ref)):
        gap = abs(lam_ref[i] - lam_ref[i - 1]) / max(abs(lam_ref[i]), 1e-6 * scale)
        if gap < cluster_rtol:
            groups[-1].append(i)
        else:
            groups.append([i])
    return groups


def subspace_angles(x, y, mass, lam_ref, cluster_rtol=1e-3, complete=None):
    """Largest principal angle per cluster between M-orthonormal bases x, y.

    ``sin(theta_max) = || (I - Y Y^T M) X ||_M`` on the cluster's columns.
    A trailing cluster that the k columns cut short (the k-th eigenvalue's
    cluster continues past k) is only compared when ``complete`` says so; its
    basis is not determined by k columns alone.
    """
    out = []
    groups = clusters(lam_ref, cluster_rtol)
    for gi, g in enumerate(groups):
        if complete is not None and gi == len(groups) - 1 and not complete:
            continue
        xs = x[:, g]
        ys = y[:, g]
        proj = ys @ (ys.T @ (mass @ xs))
        d = xs - proj
        s = np.sqrt(max(np.linalg.eigvalsh(d.T @ (mass @ d)).max(), 0.0))
        out.append((g, float(np.arcsin(min(s, 1.0)))))
    return out


def projector_probes(x, mass, lam_ref, idx, nr=4, seed=11, cluster_rtol=1e-3):
    """Basis-free samples of each cluster's eigenspace.

    For every cluster ``g`` of ``lam_ref`` (see ``clusters``) and ``nr`` seeded
    standard-normal vectors ``r``, returns ``(P_g r)[idx]`` with the
    M-orthogonal projector ``P_g = X_g X_g^T M`` onto the span of the cluster's
    M-orthonormal columns.  ``P_g`` does not depend on the signs or the basis
    chosen inside the cluster, so two solvers' results can be compared at a
    few hundred vertices: ``|P_g - P'_g| r`` is bounded by ``2 sin(theta)``
    times ``|r|_M`` for the largest principal angle ``theta`` between the
    spaces.  Output: a list of ``(columns, array (nr, len(idx)))``.
    """
    n = x.shape[0]
    r = np.random.default_rng(seed).standard_normal((n, nr))
    mr = mass @ r
    out = []
    for g in clusters(lam_ref, cluster_rtol):
        xg = x[:, g]
        p = xg @ (xg.T @ mr)
        out.append((list(g), np.asarray(p[idx]).T.copy()))
    return out


def sign_normalise(x):
    """Flip each column so that its largest-magnitude entry is positive."""
    x = np.array(x, copy=True)
    idx = np.argmax(np.abs(x), axis=0)
    s = np.sign(x[idx, np.arange(x.shape[1])])
    s[s == 0] = 1
    return x * s


def m_orthonormality_err(x, mass):
    g = x.T @ (mass @ x)
    return float(np.max(np.abs(g - np.eye(g.shape[0]))))
