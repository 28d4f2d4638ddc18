"""numpy/scipy restatement of LaPy's FEM hot path (test infrastructure only).

Every function follows the reference line by line in its arithmetic (same
numpy operations in the same order, element math in the vertex dtype, COO
entry order, ``csc_matrix`` built without ``shape=``), because the GPU
library is checked against exactly these numbers.
"""

from __future__ import annotations

import sys

import numpy as np
from scipy import sparse

_EPS = sys.float_info.epsilon


# ---------------------------------------------------------------------------
# element kernels (local values in COO order)
# ---------------------------------------------------------------------------

def _clamp(vol):
    # lapy/solver.py:158-159 (same at :272-273, :357-358, :436-437)
    vol_mean = max(0.0001 * np.mean(vol), _EPS)
    vol[vol < _EPS] = vol_mean
    return vol


def tria_local(v, t, dtype=np.float64):
    """Per-triangle COO triplets of A and B exactly as lapy/solver.py:141-187.

    Returns ``(i, j, local_a, local_b_cons, local_b_lump, i_lump)``.
    """
    dtype = np.dtype(dtype)
    t1, t2, t3 = t[:, 0], t[:, 1], t[:, 2]
    v1, v2, v3 = v[t1, :], v[t2, :], v[t3, :]
    v2mv1 = v2 - v1
    v3mv2 = v3 - v2
    v1mv3 = v1 - v3
    cr = np.cross(v3mv2, v1mv3)
    vol = 2 * np.sqrt(np.sum(cr * cr, axis=1))
    vol = _clamp(vol)
    a12 = np.sum(v3mv2 * v1mv3, axis=1) / vol
    a23 = np.sum(v1mv3 * v2mv1, axis=1) / vol
    a31 = np.sum(v2mv1 * v3mv2, axis=1) / vol
    a11 = -a12 - a31
    a22 = -a12 - a23
    a33 = -a31 - a23
    local_a = np.column_stack((a12, a12, a23, a23, a31, a31, a11, a22, a33)).reshape(-1).astype(dtype, copy=False)
    i = np.column_stack((t1, t2, t2, t3, t3, t1, t1, t2, t3)).reshape(-1)
    j = np.column_stack((t2, t1, t3, t2, t1, t3, t1, t2, t3)).reshape(-1)
    b_ii = vol / 24
    b_ij = vol / 48
    local_b = np.column_stack((b_ij, b_ij, b_ij, b_ij, b_ij, b_ij, b_ii, b_ii, b_ii)).reshape(-1).astype(dtype, copy=False)
    b_l = vol / 12
    local_bl = np.column_stack((b_l, b_l, b_l)).reshape(-1).astype(dtype, copy=False)
    il = np.column_stack((t1, t2, t3)).reshape(-1)
    return i, j, local_a, local_b, local_bl, il


def tria_aniso_local(v, t, u1, u2, aniso_mat, dtype=np.float64):
    """Per-triangle COO triplets of lapy/solver.py:249-301 (anisotropic)."""
    dtype = np.dtype(dtype)
    t1, t2, t3 = t[:, 0], t[:, 1], t[:, 2]
    v1, v2, v3 = v[t1, :], v[t2, :], v[t3, :]
    v2mv1 = v2 - v1
    v3mv2 = v3 - v2
    v1mv3 = v1 - v3
    uv2mv1 = np.column_stack((np.sum(u1 * v2mv1, axis=1), np.sum(u2 * v2mv1, axis=1)))
    uv3mv2 = np.column_stack((np.sum(u1 * v3mv2, axis=1), np.sum(u2 * v3mv2, axis=1)))
    uv1mv3 = np.column_stack((np.sum(u1 * v1mv3, axis=1), np.sum(u2 * v1mv3, axis=1)))
    cr = np.cross(v3mv2, v1mv3)
    vol = 2 * np.sqrt(np.sum(cr * cr, axis=1))
    vol = _clamp(vol)
    a12 = np.sum(uv3mv2 * aniso_mat * uv1mv3, axis=1) / vol
    a23 = np.sum(uv1mv3 * aniso_mat * uv2mv1, axis=1) / vol
    a31 = np.sum(uv2mv1 * aniso_mat * uv3mv2, axis=1) / vol
    a11 = -a12 - a31
    a22 = -a12 - a23
    a33 = -a31 - a23
    local_a = np.column_stack((a12, a12, a23, a23, a31, a31, a11, a22, a33)).reshape(-1).astype(dtype, copy=False)
    i = np.column_stack((t1, t2, t2, t3, t3, t1, t1, t2, t3)).reshape(-1)
    j = np.column_stack((t2, t1, t3, t2, t1, t3, t1, t2, t3)).reshape(-1)
    b_ii = vol / 24
    b_ij = vol / 48
    local_b = np.column_stack((b_ij, b_ij, b_ij, b_ij, b_ij, b_ij, b_ii, b_ii, b_ii)).reshape(-1).astype(dtype, copy=False)
    b_l = vol / 12
    local_bl = np.column_stack((b_l, b_l, b_l)).reshape(-1).astype(dtype, copy=False)
    il = np.column_stack((t1, t2, t3)).reshape(-1)
    return i, j, local_a, local_b, local_bl, il


def tria_mass_local(v, t, dtype=np.float64):
    """Per-triangle COO triplets of lapy/solver.py:344-376 (mass only, vol = area)."""
    dtype = np.dtype(dtype)
    t1, t2, t3 = t[:, 0], t[:, 1], t[:, 2]
    v1, v2, v3 = v[t1, :], v[t2, :], v[t3, :]
    v3mv2 = v3 - v2
    v1mv3 = v1 - v3
    cr = np.cross(v3mv2, v1mv3)
    vol = 0.5 * np.sqrt(np.sum(cr * cr, axis=1))
    vol = _clamp(vol)
    b_ii = vol / 6
    b_ij = vol / 12
    local_b = np.column_stack((b_ij, b_ij, b_ij, b_ij, b_ij, b_ij, b_ii, b_ii, b_ii)).reshape(-1).astype(dtype, copy=False)
    i = np.column_stack((t1, t2, t2, t3, t3, t1, t1, t2, t3)).reshape(-1)
    j = np.column_stack((t2, t1, t3, t2, t1, t3, t1, t2, t3)).reshape(-1)
    b_l = vol / 3
    local_bl = np.column_stack((b_l, b_l, b_l)).reshape(-1).astype(dtype, copy=False)
    il = np.column_stack((t1, t2, t3)).reshape(-1)
    return i, j, local_b, local_bl, il


def tet_local(v, t, dtype=np.float64):
    """Per-tetrahedron COO triplets of lapy/solver.py:414-532."""
    dtype = np.dtype(dtype)
    t1, t2, t3, t4 = t[:, 0], t[:, 1], t[:, 2], t[:, 3]
    v1, v2, v3, v4 = v[t1, :], v[t2, :], v[t3, :], v[t4, :]
    e1 = v2 - v1
    e2 = v3 - v2
    e3 = v1 - v3
    e4 = v4 - v1
    e5 = v4 - v2
    e6 = v4 - v3
    cr = np.cross(e1, e3)
    vol = np.abs(np.sum(e4 * cr, axis=1))
    vol = _clamp(vol)
    e11 = np.sum(e1 * e1, axis=1)
    e22 = np.sum(e2 * e2, axis=1)
    e33 = np.sum(e3 * e3, axis=1)
    e44 = np.sum(e4 * e4, axis=1)
    e55 = np.sum(e5 * e5, axis=1)
    e66 = np.sum(e6 * e6, axis=1)
    e12 = np.sum(e1 * e2, axis=1)
    e13 = np.sum(e1 * e3, axis=1)
    e14 = np.sum(e1 * e4, axis=1)
    e15 = np.sum(e1 * e5, axis=1)
    e23 = np.sum(e2 * e3, axis=1)
    e25 = np.sum(e2 * e5, axis=1)
    e26 = np.sum(e2 * e6, axis=1)
    e34 = np.sum(e3 * e4, axis=1)
    e36 = np.sum(e3 * e6, axis=1)
    a12 = (-e36 * e26 + e23 * e66) / vol
    a13 = (-e15 * e25 + e12 * e55) / vol
    a14 = (e23 * e26 - e36 * e22) / vol
    a23 = (-e14 * e34 + e13 * e44) / vol
    a24 = (e13 * e34 - e14 * e33) / vol
    a34 = (-e14 * e13 + e11 * e34) / vol
    a11 = -a12 - a13 - a14
    a22 = -a12 - a23 - a24
    a33 = -a13 - a23 - a34
    a44 = -a14 - a24 - a34
    local_a = np.column_stack((a12, a12, a23, a23, a13, a13, a14, a14, a24, a24, a34, a34, a11, a22, a33, a44)).reshape(-1)
    i = np.column_stack((t1, t2, t2, t3, t3, t1, t1, t4, t2, t4, t3, t4, t1, t2, t3, t4)).reshape(-1)
    j = np.column_stack((t2, t1, t3, t2, t1, t3, t4, t1, t4, t2, t4, t3, t1, t2, t3, t4)).reshape(-1)
    local_a = (local_a / 6.0).astype(dtype, copy=False)
    bii = vol / 60.0
    bij = vol / 120.0
    local_b = np.column_stack((bij,) * 12 + (bii,) * 4).reshape(-1).astype(dtype, copy=False)
    bl = vol / 24.0
    local_bl = np.column_stack((bl, bl, bl, bl)).reshape(-1).astype(dtype, copy=False)
    il = np.column_stack((t1, t2, t3, t4)).reshape(-1)
    return i, j, local_a, local_b, local_bl, il


# ---------------------------------------------------------------------------
# matrices as the reference builds them (scipy COO -> CSC, duplicates summed)
# ---------------------------------------------------------------------------

def fem_tria(v, t, lump=False, dtype=np.float64):
    """lapy/solver.py:106-194 -> (A, B) as ``scipy.sparse.csc_matrix``."""
    i, j, la, lb, lbl, il = tria_local(v, t, dtype)
    a = sparse.csc_matrix((la, (i, j)), dtype=dtype)
    if not lump:
        b = sparse.csc_matrix((lb, (i, j)), dtype=dtype)
    else:
        b = sparse.csc_matrix((lbl, (il, il)), dtype=dtype)
    return a, b


def fem_tria_aniso(v, t, u1, u2, aniso_mat, lump=False, dtype=np.float64):
    """lapy/solver.py:197-308."""
    i, j, la, lb, lbl, il = tria_aniso_local(v, t, u1, u2, aniso_mat, dtype)
    a = sparse.csc_matrix((la, (i, j)), dtype=dtype)
    if not lump:
        b = sparse.csc_matrix((lb, (i, j)), dtype=dtype)
    else:
        b = sparse.csc_matrix((lbl, (il, il)), dtype=dtype)
    return a, b


def fem_tria_mass(v, t, lump=False, dtype=np.float64):
    """lapy/solver.py:311-377."""
    i, j, lb, lbl, il = tria_mass_local(v, t, dtype)
    if not lump:
        return sparse.csc_matrix((lb, (i, j)), dtype=dtype)
    return sparse.csc_matrix((lbl, (il, il)), dtype=dtype)


def fem_tetra(v, t, lump=False, dtype=np.float64):
    """lapy/solver.py:380-533."""
    i, j, la, lb, lbl, il = tet_local(v, t, dtype)
    a = sparse.csc_matrix((la, (i, j)), dtype=dtype)
    if not lump:
        b = sparse.csc_matrix((lb, (i, j)), dtype=dtype)
    else:
        b = sparse.csc_matrix((lbl, (il, il)), dtype=dtype)
    return a, b


def element_order_values(pattern: sparse.csc_matrix, i, j, vals):
    """Sum COO triplets into ``pattern``'s slots in COO (element-major) order.

    scipy sums duplicates in an order fixed by an unstable ``std::sort``; this
    variant fixes the order to "ascending element, then local entry", which is
    the order the GPU library uses, so its values can be compared bit for bit.
    ``pattern`` must be symmetric (CSC arrays read as CSR: row r's columns are
    ``indices[indptr[r]:indptr[r+1]]``).
    """
    indptr = pattern.indptr
    indices = pattern.indices
    n = pattern.shape[0]
    # slot of (i, j): position of column j inside row i (rows sorted).
    rows = np.repeat(np.arange(n), np.diff(indptr))
    keys = rows.astype(np.int64) * n + indices
    q = i.astype(np.int64) * n + j.astype(np.int64)
    slot = np.searchsorted(keys, q)
    assert np.all(keys[slot] == q)
    out = np.zeros(indices.shape[0], dtype=vals.dtype)
    np.add.at(out, slot, vals)
    return out


# ---------------------------------------------------------------------------
# mesh helpers used by the solves
# ---------------------------------------------------------------------------

def avg_edge_length(v, t):
    """lapy/tria_mesh.py:735-748 and lapy/tet_mesh.py:182-195.

    The edge set is the upper triangle of the element adjacency, built as in
    tria_mesh.py:486-509 / tet_mesh.py:103-122.
    """
    nodes = t.shape[1]
    if nodes == 3:
        t0, t1, t2 = t[:, 0], t[:, 1], t[:, 2]
        i = np.column_stack((t0, t1, t1, t2, t2, t0)).reshape(-1)
        j = np.column_stack((t1, t0, t2, t1, t0, t2)).reshape(-1)
        n = v.shape[0]
        adj = sparse.csc_matrix((np.ones(i.shape), (i, j)), shape=(n, n))
    else:
        t1, t2, t3, t4 = t[:, 0], t[:, 1], t[:, 2], t[:, 3]
        i = np.hstack((t1, t2, t2, t3, t3, t1, t1, t2, t3, t4, t4, t4))
        j = np.hstack((t2, t1, t3, t2, t1, t3, t4, t4, t4, t1, t2, t3))
        adj = sparse.csc_matrix((np.ones(i.shape, dtype=int), (i, j)))
    triadj = sparse.triu(adj, 1, format="coo")
    edgelens = np.sqrt(((v[triadj.row, :] - v[triadj.col, :]) ** 2).sum(1))
    return edgelens.mean()


# ---------------------------------------------------------------------------
# solves
# ---------------------------------------------------------------------------

def eigs(a, b, k=10, sigma=-0.01, v0=None):
    """lapy/solver.py:667-716: shift-invert ``eigsh`` with an ``splu`` OPinv."""
    from scipy.sparse.linalg import LinearOperator, eigsh, splu

    lu = splu(a - sigma * b)
    op_inv = LinearOperator(matvec=lu.solve, shape=a.shape, dtype=a.dtype)
    return eigsh(a, k, b, sigma=sigma, OPinv=op_inv, v0=v0)


def poisson(stiffness, mass, h=0.0, dtup=(), ntup=()):
    """lapy/solver.py:718-889 (splu path)."""
    from scipy.sparse.linalg import splu

    dim = stiffness.shape[0]
    dtype = stiffness.dtype
    if stiffness.shape != mass.shape or stiffness.shape[1] != dim:
        raise ValueError("Error: Square input matrices should have same number of rows and columns.")
    if np.isscalar(h):
        h = np.full((dim, 1), h, dtype=dtype)
    else:
        h = np.asarray(h, dtype=dtype)
        if h.ndim == 1:
            if h.size != dim:
                raise ValueError("h should be either scalar or column vector with row num of A")
            h = h[:, np.newaxis]
        elif h.ndim == 2:
            if h.shape[0] != dim:
                raise ValueError("h should be either scalar or array with first dim matching A")
        else:
            raise ValueError("h should be either scalar or 1-D/2-D array")
    n_rhs = h.shape[1]
    scalar_rhs = n_rhs == 1
    didx = []
    dvec = []
    if dtup:
        if len(dtup) != 2:
            raise ValueError("dtup should contain index and data arrays")
        didx, ddat = dtup[0], dtup[1]
        if np.unique(didx).size != len(didx):
            raise ValueError("dtup indices need to be unique")
        if not (len(didx) > 0 and len(didx) == len(ddat)):
            raise ValueError("dtup should contain index and data arrays (same lengths > 0)")
        dvec = sparse.csc_matrix(
            (np.tile(ddat, n_rhs), (np.tile(didx, n_rhs), np.repeat(np.arange(n_rhs), len(didx)))),
            (dim, n_rhs), dtype=dtype)
    nvec = 0
    if ntup:
        if len(ntup) != 2:
            raise ValueError("ntup should contain index and data arrays")
        nidx, ndat = ntup[0], ntup[1]
        if not (len(nidx) > 0 and len(nidx) == len(ndat)):
            raise ValueError("ntup should contain index and data arrays (same lengths > 0)")
        nvec = sparse.csc_matrix(
            (np.tile(ndat, n_rhs), (np.tile(nidx, n_rhs), np.repeat(np.arange(n_rhs), len(nidx)))),
            (dim, n_rhs), dtype=dtype)
    mass = mass.astype(dtype, copy=False)
    bvec = mass * (h - nvec)
    if len(didx) > 0:
        bvec = bvec - stiffness * dvec
    mask = []
    if len(didx) > 0:
        mask = np.full(dim, True, dtype=bool)
        mask[didx] = False
        bvec = bvec[mask]
        amat = stiffness.tocsc()[:, mask].tocsr()
        amat = amat[mask, :].tocsc()
    else:
        amat = stiffness
    x = splu(amat).solve(np.asarray(bvec).astype(dtype))
    if len(didx) > 0:
        xfull = np.zeros((dim, n_rhs), dtype=dtype)
        xfull[mask, :] = x
        xfull[didx, :] = dvec[didx, :].toarray()
        x = xfull
    return np.squeeze(np.array(x)) if scalar_rhs else np.array(x)


def diffusion_matrix_rhs(v, t, vids, m=1.0, stiffness=None, mass=None):
    """Heat matrix and right-hand side of lapy/heat.py:185-214."""
    nv = len(v)
    if isinstance(vids, list) and len(vids) > 0 and isinstance(vids[0], (list, np.ndarray)):
        scalar_input = False
        vids_list = [np.asarray(x, dtype=int).ravel() for x in vids]
    else:
        scalar_input = True
        vids_list = [np.asarray(vids, dtype=int).ravel()]
    for x in vids_list:
        if np.any(x < 0) or np.any(x >= nv):
            raise ValueError("vids contains out-of-range vertex indices")
    if stiffness is None:
        if t.shape[1] == 3:
            stiffness, mass = fem_tria(v, t, lump=True)
        else:
            stiffness, mass = fem_tetra(v, t, lump=True)
    tt = m * avg_edge_length(v, t) ** 2
    hmat = mass + tt * stiffness
    b0 = np.zeros((nv, len(vids_list)))
    for kk, x in enumerate(vids_list):
        b0[x, kk] = 1.0
    return hmat, b0, scalar_input, tt


def diffusion(v, t, vids, m=1.0):
    """lapy/heat.py:114-232 (splu path)."""
    from scipy.sparse.linalg import splu

    hmat, b0, scalar_input, _ = diffusion_matrix_rhs(v, t, vids, m)
    vfunc = splu(hmat).solve(b0.astype(np.float64))
    return vfunc[:, 0] if scalar_input else vfunc
