"""CPU restatement of the reference's principal curvatures (test
infrastructure: the GPU curvature, lh_curvature / lh_curvature_tria, is
checked against these functions, and they are pinned by geometry in
tests/test_curvature.py).

Follows lapy/tria_mesh.py:1078-1219 (curvature), :1221-1279
(curvature_tria), :750-817 (triangle and vertex normals), :1019-1076
(inner edges) and :1696-1792 (the smoothing).  The estimator is the
edge-based curvature tensor of Alliez, Cohen-Steiner, Devillers, Levy and
Desbrun ("Anisotropic Polygonal Remeshing", ACM TOG 2003): each inner edge
contributes its signed dihedral angle times its (mean-normalised) length
times ``e e^T`` for its unit direction ``e``; a vertex averages the
contributions of its edges, the six tensor entries are smoothed by neighbour
averaging, and the per-vertex 3x3 eigenproblem (numpy's eigh, as the
reference) gives the normal (the eigenvector most aligned with the vertex
normal) and the two principal directions.
"""

import sys

import numpy as np
from scipy import sparse


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
    products, i.e. area-weighted triangle normals (``tria_mesh.py:775-817``).
    Raises ``ValueError`` for a mesh that is not consistently oriented."""
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


def inner_edges(v, t):
    """Inner edges of an oriented mesh (``tria_mesh.py:1019-1076``):
    ``vids`` (E, 2) with ``vids[:,0] < vids[:,1]``, sorted lexicographically,
    and ``tids`` (E, 2): the triangle holding the half edge ``vids[:,0] ->
    vids[:,1]`` and the triangle holding the opposite half edge. Boundary
    edges (one triangle) are left out."""
    n = v.shape[0]
    _check_oriented(t, n)
    nt = t.shape[0]
    a = t.reshape(-1)
    b = t[:, [1, 2, 0]].reshape(-1)
    tri = np.repeat(np.arange(nt), 3)
    fwd = a < b
    key_f = a[fwd].astype(np.int64) * n + b[fwd]
    key_b = b[~fwd].astype(np.int64) * n + a[~fwd]
    keys, i_f, i_b = np.intersect1d(key_f, key_b, assume_unique=True, return_indices=True)
    vids = np.stack([keys // n, keys % n], axis=1)
    tids = np.stack([tri[fwd][i_f], tri[~fwd][i_b]], axis=1)
    return vids, tids


def _smoothing_matrix(v, t):
    """Row-stochastic neighbour average (``tria_mesh.py:1696-1715``): row i
    holds ``1 / deg_i`` at each of its ``deg_i`` distinct edge neighbours. The
    edge adjacency is binarised first (``tria_mesh.py:1707``), so an inner
    edge (two triangles) and a boundary edge (one) weigh the same. The
    reference weights both by the vertex area, which cancels; a vertex of zero
    area keeps a zero row, as there."""
    n = v.shape[0]
    i = t[:, [0, 1, 1, 2, 2, 0]].reshape(-1)
    j = t[:, [1, 0, 2, 1, 0, 2]].reshape(-1)
    adj = sparse.csr_matrix((np.ones(i.size), (i, j)), shape=(n, n))
    adj.sum_duplicates()
    adj.data[:] = 1.0
    v0, v1, v2 = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    cr = np.cross(v1 - v0, v2 - v0)
    area = 0.5 * np.sqrt(np.sum(cr * cr, axis=1))
    varea = np.bincount(t.reshape(-1), np.repeat(area, 3), minlength=n) / 3.0
    deg = np.asarray(adj.sum(axis=1)).ravel()
    scale = np.where((varea * deg) != 0, 1.0 / np.where(deg == 0, 1.0, deg), 0.0)
    return sparse.diags(scale) @ adj


def vertex_tensors(v, t, smoothit=3):
    """The smoothed per-vertex curvature tensors (V, 3, 3) whose eigenproblems
    give the principal curvatures (``tria_mesh.py:1078-1190``)."""
    v = np.asarray(v)
    t = np.asarray(t)
    n = v.shape[0]
    vids, tids = inner_edges(v, t)
    tn = tria_normals(v, t)
    n0, n1 = tn[tids[:, 0]], tn[tids[:, 1]]
    angle = np.arccos(np.clip(np.sum(n0 * n1, axis=1), -1.0, 1.0))
    evec = v[vids[:, 1]] - v[vids[:, 0]]
    elen = np.sqrt(np.sum(evec * evec, axis=1))
    # signed dihedral angle: negative where the surface is convex (outward normals)
    angle = angle * -np.sign(np.sum(np.cross(n0, n1) * evec, axis=1))
    elen[elen < sys.float_info.epsilon] = 1
    e = evec / elen[:, None]
    elen = elen / np.mean(elen)
    # upper triangle of e e^T: xx, xy, xz, yy, yz, zz
    ee = np.stack([e[:, 0] * e[:, 0], e[:, 0] * e[:, 1], e[:, 0] * e[:, 2],
                   e[:, 1] * e[:, 1], e[:, 1] * e[:, 2], e[:, 2] * e[:, 2]], axis=1)
    ee = ee * (angle * elen)[:, None]
    ends = np.concatenate([vids[:, 0], vids[:, 1]])
    vv = np.stack([np.bincount(ends, np.concatenate([ee[:, c], ee[:, c]]), minlength=n)
                   for c in range(6)], axis=1)
    deg = np.bincount(ends, minlength=n).astype(np.float64)
    deg[deg == 0] = 1
    vv = vv / deg[:, None]
    mat = _smoothing_matrix(v, t)
    for _ in range(smoothit):
        vv = mat @ vv
    return vv[:, [0, 1, 2, 1, 3, 4, 2, 4, 5]].reshape(n, 3, 3)


def curvature(v, t, smoothit=3):
    """Per-vertex principal curvatures (``tria_mesh.py:1078-1219``).

    Returns ``u_min, u_max`` (V, 3) unit principal directions, ``c_min, c_max``
    (V,) with ``c_min <= c_max``, ``c_mean = (c_min + c_max) / 2`` and
    ``c_gauss = c_min * c_max`` (taken before the two are ordered, as in the
    reference), and ``normals`` (V, 3), the tensor's normal eigenvector turned
    towards the vertex normal. ``u_max`` is oriented so that
    ``u_min x u_max`` points along ``normals``.
    """
    v = np.asarray(v)
    t = np.asarray(t)
    tens = vertex_tensors(v, t, smoothit)
    evals, evecs = np.linalg.eigh(tens)
    vn = vertex_normals(v, t)
    # the eigenvector most aligned with the vertex normal first (not the
    # smallest |eigenvalue|, which is ambiguous on flat regions)
    order = np.argsort(-np.abs(np.einsum("vrk,vr->vk", evecs, vn)), axis=1)
    evals = np.take_along_axis(evals, order, axis=1)
    evecs = np.take_along_axis(evecs, order[:, None, :], axis=2)
    normals = evecs[:, :, 0].copy()
    u_max = evecs[:, :, 1].copy()
    u_min = evecs[:, :, 2].copy()
    c_max = evals[:, 2].copy()
    c_min = evals[:, 1].copy()
    c_mean = (c_min + c_max) / 2.0
    c_gauss = c_min * c_max
    swap = c_min > c_max
    c_min[swap], c_max[swap] = c_max[swap], c_min[swap]
    u_min[swap], u_max[swap] = u_max[swap], u_min[swap]
    normals = normals * np.sign(np.sum(normals * vn, axis=1))[:, None]
    flip = np.sum(np.cross(u_min, u_max) * normals, axis=1) < 0
    u_max[flip] = -u_max[flip]
    return u_min, u_max, c_min, c_max, c_mean, c_gauss, normals


def curvature_tria(v, t, smoothit=3):
    """Per-triangle principal directions and curvatures
    (``tria_mesh.py:1221-1279``): the vertex values averaged over each
    triangle's corners; ``u_min`` projected into the triangle plane and
    normalised, ``u_max = normal x u_min``. Returns ``u_min, u_max`` (F, 3)
    and ``c_min, c_max`` (F,)."""
    v = np.asarray(v)
    t = np.asarray(t)
    u_min, _u_max, c_min, c_max, _m, _g, _n = curvature(v, t, smoothit)
    tumin = np.sum((u_min / 3.0)[t], axis=1)
    tcmin = np.sum((c_min / 3.0)[t], axis=1)
    tcmax = np.sum((c_max / 3.0)[t], axis=1)
    tn = tria_normals(v, t)
    tumin = tumin - tn * np.sum(tn * tumin, axis=1)[:, None]
    tumin = tumin / np.maximum(np.sqrt(np.sum(tumin * tumin, axis=1)), 1e-8)[:, None]
    tumax = np.cross(tn, tumin)
    return tumin, tumax, tcmin, tcmax
