"""Gradient, divergence and heat-method geodesic distance on triangle and
tetrahedral meshes (lapy/diffgeo.py:26-164, 222-387, 846-1006), and mean
curvature flow of triangle meshes (lapy/diffgeo.py:523-607), on the GPU.

Same names, arguments, dispatch and return shapes as the reference.  The
element work (gradient, normalisation, divergence terms) and the vertex sums
run in ``lh_diffgeo``; the Poisson solve of ``compute_geodesic_f`` is the
library's PCG on the lumped Solver's stiffness with the mass replaced by the
identity, as the reference's ``fem.mass = sparse.eye(n)`` does.
"""

from __future__ import annotations

import numpy as np

from . import _lib
from .mesh import device_mesh
from .solver import Solver


def _kind(geom):
    name = type(geom).__name__
    if name in ("TriaMesh", "TetMesh"):
        return name
    raise ValueError('Geometry type "' + name + '" unknown')


def _dim(geom):
    t = np.asarray(geom.t)
    return int(t.max()) + 1 if t.size else 0


def compute_gradient(geom, vfunc: np.ndarray) -> np.ndarray:
    """Per-element gradient of a vertex function (lapy/diffgeo.py:26-54,
    222-300, 846-918): ``(n_elements, 3)`` for ``(n_vertices,)`` input,
    ``(n_elements, n_functions, 3)`` for ``(n_vertices, n_functions)``."""
    _kind(geom)
    vfunc = np.asarray(vfunc)
    scalar = vfunc.ndim == 1
    f2 = vfunc[:, None] if scalar else vfunc
    dm = device_mesh(geom)
    if f2.shape[0] != dm.nv:
        raise ValueError("vfunc must have one row per vertex")
    g = dm.diffgeo(_lib.LH_DG_GRADIENT, f2, f2.shape[1])
    return g[:, 0, :] if scalar else g


def compute_divergence(geom, tfunc: np.ndarray) -> np.ndarray:
    """Integrated divergence at the vertices of a per-element vector field
    (lapy/diffgeo.py:57-85, 303-387, 921-1006): ``(n,)`` for
    ``(n_elements, 3)`` input (n = max(t)+1, the inferred shape of the
    reference's ``csc_matrix``), ``(n_vertices, n_functions)`` for
    ``(n_elements, n_functions, 3)``."""
    _kind(geom)
    tfunc = np.asarray(tfunc)
    scalar = tfunc.ndim == 2
    f3 = tfunc[:, None, :] if scalar else tfunc
    dm = device_mesh(geom)
    if f3.shape[0] != dm.ne or f3.shape[-1] != 3:
        raise ValueError("tfunc must have shape (n_elements, 3) or (n_elements, n_functions, 3)")
    d = dm.diffgeo(_lib.LH_DG_DIVERGENCE, f3, f3.shape[1])
    if scalar:
        return d[:_dim(geom), 0].astype(tfunc.dtype, copy=False)
    return d


def compute_geodesic_f(geom, vfunc: np.ndarray, use_cholmod: bool = False) -> np.ndarray:
    """Function whose gradient is the normalised gradient of ``vfunc``
    (geodesic distance by the heat method, lapy/diffgeo.py:117-164): unit
    gradient field, its divergence, ``Solver(lump=True).poisson`` with the
    mass replaced by the identity, shifted so that its minimum is 0."""
    _kind(geom)
    vfunc = np.asarray(vfunc)
    f2 = vfunc[:, None] if vfunc.ndim == 1 else vfunc
    fem = Solver(geom, lump=True, use_cholmod=use_cholmod)
    dm = device_mesh(geom, fem._ctx)
    if f2.shape[0] != dm.nv:
        raise ValueError("vfunc must have one row per vertex")
    divf = dm.diffgeo(_lib.LH_DG_GEODESIC_RHS, f2, f2.shape[1])
    a_dev = fem.device_matrices()[0]
    if divf.shape[0] != a_dev.n:
        raise ValueError("h should be either scalar or array with first dim matching A")
    # Solver.poisson(divf) with B = I and no boundary data is the solve A x = divf; A is the singular
    # Neumann Laplacian and divf sums to zero per element, so CG converges; the constant is removed below
    x, st = _lib.pcg(fem._ctx, a_dev, None, 0.0, 1.0, divf, rtol=1e-12)
    fem.last_stats = st
    vf = x[:, 0] if x.shape[1] == 1 else x
    return vf - vf.min(axis=0)


def _normalized(v, t):
    """TriaMesh.normalize_ (lapy/tria_mesh.py:990-1017, 1390-1405): centroid
    weighted by triangle area to the origin, unit total area.  An O(F) host
    reduction per flow step, as in the reference's mesh class."""
    v0, v1, v2 = v[t[:, 0], :], v[t[:, 1], :], v[t[:, 2], :]
    cr = np.cross(v2 - v1, v0 - v2)
    areas = 0.5 * np.sqrt(np.sum(cr * cr, axis=1))
    totalarea = areas.sum()
    if totalarea <= 0:
        raise ValueError("Mesh surface area must be positive to normalize.")
    centroid = np.sum((1.0 / 3.0) * (v0 + v1 + v2) * (areas / totalarea)[:, np.newaxis], axis=0)
    return (1.0 / np.sqrt(totalarea)) * (v - centroid)


def tria_mean_curvature_flow(tria, max_iter: int = 30, stop_eps: float = 1e-13, step: float = 1.0,
                             use_cholmod: bool = False):
    """Non-singular mean curvature flow (Kazhdan 2012), lapy/diffgeo.py:523-607.

    The lumped stiffness of the normalised input stays fixed; every step
    assembles the lumped mass of the current vertices on the GPU, solves
    ``(M + step*A) v = M v`` for the three coordinate columns by PCG (warm
    started from the previous vertices, relative residual 1e-12, where the
    reference calls spsolve), renormalises, and stops once
    ``trace((dv' M dv)**2) < stop_eps``.  Returns a new ``TriaMesh``."""
    from .mesh import TriaMesh
    from .solver import import_optional_dependency

    if use_cholmod:
        import_optional_dependency("sksparse", raise_error=True)
    t = np.asarray(tria.t)
    v = _normalized(np.asarray(tria.v, dtype=np.float64), t)
    ctx = _lib.default_context()
    dm0 = _lib.DeviceMesh(ctx, v, t)
    a_mat, _ = dm0.assemble(_lib.LH_TRIA, True, np.float64)
    if a_mat.n != v.shape[0]:
        raise ValueError("mean curvature flow needs every vertex in a triangle (no free vertices)")
    for _ in range(max_iter):
        vlast = v
        dm = _lib.DeviceMesh(ctx, v, t)
        _, mass = dm.assemble(_lib.LH_TRIA_MASS, True, np.float64)
        mass_v = mass.matmul(v)
        v, _st = _lib.pcg(ctx, a_mat, mass, 1.0, float(step), mass_v, x0=vlast, rtol=1e-12)
        v = _normalized(v, t)
        dv = v - vlast
        diff = np.trace(np.square(dv.T @ mass.matmul(dv)))
        if diff < stop_eps:
            break
    return TriaMesh(v, t)
