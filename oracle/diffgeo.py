"""numpy/scipy restatement of LaPy's gradient, divergence, heat-method
geodesic distance and mean curvature flow (test infrastructure only).

Same numpy operations in the same order as the reference, geometry in the
vertex dtype, per-function arithmetic in float64, so the GPU library can be
checked against these numbers.  Only tests, the smoke check and the CPU
baseline of the benchmark import this module.
"""

from __future__ import annotations

import sys

import numpy as np
from scipy import sparse

from . import fem


def _edges(v, t):
    v0 = v[t[:, 0], :]
    v1 = v[t[:, 1], :]
    v2 = v[t[:, 2], :]
    return v1 - v0, v2 - v1, v0 - v2  # e2, e0, e1


def tria_gradient(v, t, vfunc):
    """lapy/diffgeo.py:222-300: per-triangle gradient of a vertex function,
    ``(F, 3)`` for ``(n,)`` input, ``(F, nf, 3)`` for ``(n, nf)``."""
    vfunc = np.asarray(vfunc)
    e2, e0, e1 = _edges(v, t)
    n = np.cross(e2, -e1)
    ln = np.sqrt(np.sum(n * n, axis=1))
    ln[ln < sys.float_info.epsilon] = 1
    lni = np.divide(1.0, ln)[:, np.newaxis]
    n *= lni
    if vfunc.ndim == 1:
        c0 = vfunc[t[:, 0], np.newaxis] * e0
        c1 = vfunc[t[:, 1], np.newaxis] * e1
        c2 = vfunc[t[:, 2], np.newaxis] * e2
        return lni * np.cross(n, (c0 + c1 + c2))
    f0, f1, f2 = vfunc[t[:, 0], :], vfunc[t[:, 1], :], vfunc[t[:, 2], :]
    c = (f0[:, :, None] * e0[:, None, :] + f1[:, :, None] * e1[:, None, :]
         + f2[:, :, None] * e2[:, None, :])
    return lni[:, np.newaxis, :] * np.cross(n[:, np.newaxis, :], c)


def tria_divergence_terms(v, t, tfunc):
    """lapy/diffgeo.py:346-375: the three per-triangle terms x0, x1, x2
    (each ``(F,)`` or ``(F, nf)``) before the vertex sum."""
    tfunc = np.asarray(tfunc)
    e2, e0, e1 = _edges(v, t)
    n = np.cross(e2, -e1)
    ln = np.sqrt(np.sum(n * n, axis=1))
    ln[ln < sys.float_info.epsilon] = 1
    cot0 = (e2 * (-e1)).sum(1) / ln
    cot1 = (e0 * (-e2)).sum(1) / ln
    cot2 = (e1 * (-e0)).sum(1) / ln
    c0 = cot0[:, np.newaxis] * e0
    c1 = cot1[:, np.newaxis] * e1
    c2 = cot2[:, np.newaxis] * e2
    if tfunc.ndim == 2:
        return (((c2 - c1) * tfunc).sum(1), ((c0 - c2) * tfunc).sum(1), ((c1 - c0) * tfunc).sum(1))
    return (((c2 - c1)[:, None, :] * tfunc).sum(-1), ((c0 - c2)[:, None, :] * tfunc).sum(-1),
            ((c1 - c0)[:, None, :] * tfunc).sum(-1))


def tria_divergence(v, t, tfunc):
    """lapy/diffgeo.py:303-387: integrated divergence at the vertices."""
    tfunc = np.asarray(tfunc)
    x0, x1, x2 = tria_divergence_terms(v, t, tfunc)
    if tfunc.ndim == 2:
        i = np.column_stack((t[:, 0], t[:, 1], t[:, 2])).reshape(-1)
        j = np.zeros((3 * len(t), 1), dtype=int).reshape(-1)
        dat = np.column_stack((x0, x1, x2)).reshape(-1)
        return np.squeeze(np.asarray(0.5 * sparse.csc_matrix((dat, (i, j))).todense(), dtype=tfunc.dtype))
    n_t = len(t)
    rows = np.concatenate([t[:, 0], t[:, 1], t[:, 2]])
    s = sparse.csr_matrix((np.ones(3 * n_t), (rows, np.arange(3 * n_t))), shape=(v.shape[0], 3 * n_t))
    return 0.5 * s.dot(np.vstack([x0, x1, x2]))


def tet_gradient(v, t, vfunc):
    """lapy/diffgeo.py:846-918: per-tetrahedron gradient, ``(T, 3)`` or ``(T, nf, 3)``."""
    vfunc = np.asarray(vfunc)
    v0, v1, v2, v3 = (v[t[:, c], :] for c in range(4))
    e0, e2, e3, e4, e5 = v1 - v0, v0 - v2, v3 - v0, v3 - v1, v3 - v2
    cr = np.cross(e0, e2)
    vol = np.abs(np.sum(e3 * cr, axis=1))
    vol[vol < sys.float_info.epsilon] = 1
    voli = np.divide(1.0, vol)[:, np.newaxis]
    cr25, cr34, cr2e0 = np.cross(e2, e5), np.cross(e3, e4), np.cross(-e2, e0)
    if vfunc.ndim == 1:
        c1 = (vfunc[t[:, 1], np.newaxis] - vfunc[t[:, 0], np.newaxis]) * cr25
        c2 = (vfunc[t[:, 2], np.newaxis] - vfunc[t[:, 0], np.newaxis]) * cr34
        c3 = (vfunc[t[:, 3], np.newaxis] - vfunc[t[:, 0], np.newaxis]) * cr2e0
        return voli * (c1 + c2 + c3)
    d1, d2, d3 = (vfunc[t[:, c], :] - vfunc[t[:, 0], :] for c in (1, 2, 3))
    c = (d1[:, :, None] * cr25[:, None, :] + d2[:, :, None] * cr34[:, None, :] + d3[:, :, None] * cr2e0[:, None, :])
    return voli[:, np.newaxis, :] * c


def tet_divergence(v, t, tfunc):
    """lapy/diffgeo.py:921-1006: integrated divergence at the vertices,
    -1/6 of the field's flux through the faces opposite each vertex."""
    tfunc = np.asarray(tfunc)
    v0, v1, v2, v3 = (v[t[:, c], :] for c in range(4))
    e0, e1, e2, e3, e4 = v1 - v0, v2 - v1, v2 - v0, v3 - v0, v3 - v1
    ns = (np.cross(e1, e4), np.cross(e3, e2), np.cross(e0, e3), np.cross(e2, e0))
    n_t = len(t)
    if tfunc.ndim == 2:
        xs = [(n * tfunc).sum(1) for n in ns]
        i = np.column_stack((t[:, 0], t[:, 1], t[:, 2], t[:, 3])).reshape(-1)
        j = np.zeros((4 * n_t, 1), dtype=int).reshape(-1)
        dat = np.column_stack(xs).reshape(-1)
        return -np.squeeze(np.asarray((1.0 / 6.0) * sparse.csc_matrix((dat, (i, j))).todense(), dtype=tfunc.dtype))
    xs = [(n[:, np.newaxis, :] * tfunc).sum(-1) for n in ns]
    rows = np.concatenate([t[:, 0], t[:, 1], t[:, 2], t[:, 3]])
    s = sparse.csr_matrix((np.ones(4 * n_t), (rows, np.arange(4 * n_t))), shape=(v.shape[0], 4 * n_t))
    return -(1.0 / 6.0) * s.dot(np.vstack(xs))


def gradient(v, t, vfunc):
    """lapy/diffgeo.py:26-54: dispatch on nodes per element."""
    return tet_gradient(v, t, vfunc) if t.shape[1] == 4 else tria_gradient(v, t, vfunc)


def divergence(v, t, tfunc):
    """lapy/diffgeo.py:57-85."""
    return tet_divergence(v, t, tfunc) if t.shape[1] == 4 else tria_divergence(v, t, tfunc)


def normalized_gradient(gradf):
    """lapy/diffgeo.py:148-162: unit gradient, 0 where the gradient is 0."""
    if gradf.ndim == 2:
        g = gradf / np.sqrt((gradf ** 2).sum(1))[:, np.newaxis]
    else:
        g = gradf / np.sqrt((gradf ** 2).sum(-1))[:, :, np.newaxis]
    return np.nan_to_num(g)


def geodesic_f(v, t, vfunc):
    """lapy/diffgeo.py:117-164 (triangles or tetrahedra): gradient, normalise,
    divergence, then ``Solver(lump=True).poisson`` with the mass replaced by
    the identity (an splu solve of the stiffness), shifted to minimum 0."""
    vfunc = np.asarray(vfunc)
    with np.errstate(invalid="ignore", divide="ignore"):
        gradnorm = normalized_gradient(gradient(v, t, vfunc))
    divf = divergence(v, t, gradnorm)
    a, _ = (fem.fem_tetra if t.shape[1] == 4 else fem.fem_tria)(v, t, lump=True)
    vf = fem.poisson(a, sparse.eye(a.shape[0], dtype=a.dtype), divf)
    return vf - vf.min(axis=0)


def normalize(v, t):
    """lapy/tria_mesh.py:990-1017 (area-weighted centroid, total area) and
    :1390-1405 (normalize_): translate the centroid to the origin, scale to
    unit surface area."""
    v0, v1, v2 = v[t[:, 0], :], v[t[:, 1], :], v[t[:, 2], :]
    cr = np.cross(v2 - v1, v0 - v2)
    areas = 0.5 * np.sqrt(np.sum(cr * cr, axis=1))
    totalarea = areas.sum()
    areas = areas / totalarea
    centers = (1.0 / 3.0) * (v0 + v1 + v2)
    centroid = np.sum(centers * areas[:, np.newaxis], axis=0)
    if totalarea <= 0:
        raise ValueError("Mesh surface area must be positive to normalize.")
    return (1.0 / np.sqrt(totalarea)) * (v - centroid)


def mean_curvature_flow(v, t, max_iter=30, stop_eps=1e-13, step=1.0):
    """lapy/diffgeo.py:523-607: fixed lumped stiffness of the normalised
    mesh; each step solves (M + step*A) v = M v with the current lumped mass
    (spsolve), renormalises, and stops when trace((dv' M dv)**2) < stop_eps.
    Returns (vertices, steps taken)."""
    from scipy.sparse.linalg import spsolve

    v = normalize(np.asarray(v), t)
    a_mat, _ = fem.fem_tria(v, t, lump=True)
    x = 0
    for x in range(max_iter):
        vlast = v
        mass = fem.fem_tria_mass(v, t, lump=True)
        mass_v = mass.dot(v)
        v = spsolve(mass + step * a_mat, mass_v)
        v = normalize(v, t)
        dv = v - vlast
        diff = np.trace(np.square(np.matmul(np.transpose(dv), mass.dot(dv))))
        if diff < stop_eps:
            break
    return v, x + 1
