"""Principal curvatures of a triangle mesh on the GPU: the input of the
anisotropic Laplace-Beltrami assembly (``Solver(aniso=...)``).

``curvature`` and ``curvature_tria`` replace lapy/tria_mesh.py:1078-1219 and
:1221-1279 with the lh_curvature / lh_curvature_tria kernels (edge terms,
vertex tensors, smoothing, per-vertex 3x3 eigenproblems and the per-triangle
projection all on the device).  Two differences to the reference, both where
it is not determined: an eigenvector's sign is chosen so that its
largest-magnitude component is positive (the reference keeps LAPACK's), and
``curvature_tria`` turns each corner's ``u_min`` to agree with the first
corner's before averaging (the reference averages the corners as LAPACK left
them).  The anisotropic assembly does not depend on the sign of a triangle's
direction.

``tria_normals`` and ``vertex_normals`` are the host mesh utilities of
lapy/tria_mesh.py:750-817.
"""

from __future__ import annotations

import sys

import numpy as np

from . import _lib


def _check_oriented(t, n):
    he = t[:, [0, 1, 1, 2, 2, 0]].reshape(-1, 2)
    key = he[:, 0].astype(np.int64) * n + he[:, 1]
    if np.unique(key).size != key.size:
        raise ValueError("Error: Can only compute edge information for oriented meshes!")


def tria_normals(v, t):
    """Unit triangle normals ``(v1 - v0) x (v2 - v0)`` (``tria_mesh.py:750-773``);
    a zero-length normal is left unnormalised (zero)."""
    v0, v1, v2 = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    nrm = np.cross(v1 - v0, v2 - v0)
    ln = np.sqrt(np.sum(nrm * nrm, axis=1))
    ln[ln < sys.float_info.epsilon] = 1
    return nrm / ln[:, None]


def vertex_normals(v, t):
    """Unit vertex normals: the sum of the incident triangles' corner cross
    products (``tria_mesh.py:775-817``). Raises ``ValueError`` for a mesh that
    is not consistently oriented."""
    n = v.shape[0]
    _check_oriented(t, n)
    v0, v1, v2 = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    corner = (np.cross(v1 - v0, v2 - v0), np.cross(v2 - v1, v0 - v1), np.cross(v0 - v2, v1 - v2))
    idx = np.concatenate([t[:, 0], t[:, 1], t[:, 2]])
    w = np.concatenate(corner)
    acc = np.stack([np.bincount(idx, w[:, c], minlength=n) for c in range(3)], axis=1)
    ln = np.sqrt(np.sum(acc * acc, axis=1))
    ln[ln < sys.float_info.epsilon] = 1
    return acc / ln[:, None]


def _device_mesh(v, t, dm):
    v = np.asarray(v)
    t = np.asarray(t)
    if t.ndim != 2 or t.shape[1] != 3:
        raise ValueError("curvature needs a triangle mesh")
    _check_oriented(t, v.shape[0])
    if dm is not None:
        return dm
    if v.dtype not in (np.float32, np.float64):
        v = v.astype(np.float64)
    return _lib.DeviceMesh(_lib.default_context(), v, t)


def curvature(v, t, smoothit: int = 3, dm=None):
    """Per-vertex principal curvatures (``tria_mesh.py:1078-1219``) on the GPU.

    Returns ``u_min, u_max`` (V, 3) unit principal directions, ``c_min, c_max``
    (V,) with ``c_min <= c_max``, ``c_mean = (c_min + c_max) / 2``,
    ``c_gauss = c_min * c_max`` (taken before the two are ordered, as in the
    reference) and ``normals`` (V, 3), turned towards the vertex normals;
    ``u_max`` is oriented so that ``u_min x u_max`` points along ``normals``.
    ``dm``: an existing device mesh of (v, t).
    """
    dm = _device_mesh(v, t, dm)
    return _lib.curvature(dm.ctx, dm, int(smoothit))


def curvature_tria(v, t, smoothit: int = 3, dm=None):
    """Per-triangle principal directions and curvatures
    (``tria_mesh.py:1221-1279``) on the GPU: the vertex values averaged over
    each triangle's corners, ``u_min`` projected into the triangle plane and
    normalised, ``u_max = normal x u_min``. Returns ``u_min, u_max`` (F, 3)
    and ``c_min, c_max`` (F,)."""
    dm = _device_mesh(v, t, dm)
    return _lib.curvature_tria(dm.ctx, dm, int(smoothit))
