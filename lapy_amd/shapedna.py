"""Drop-in ``shapedna.compute_shapedna`` (lapy/shapedna.py:96-166)."""

from __future__ import annotations

import numpy as np

from . import _lib
from .mesh import device_mesh
from .solver import Solver


def compute_shapedna(geom, k: int = 50, lump: bool = False, aniso=None, aniso_smooth: int = 10,
                     use_cholmod: bool = False, **eig_opts) -> dict:
    """Eigenvalues / eigenvectors of the Laplace-Beltrami operator as a dict
    with the reference's keys."""
    fem = Solver(geom, lump=lump, aniso=aniso, aniso_smooth=aniso_smooth, use_cholmod=use_cholmod)
    evals, evecs = fem.eigs(k=k, **eig_opts)
    return _shapedna_dict(geom, k, evals, evecs)


def _shapedna_dict(geom, k, evals, evecs) -> dict:
    ev = dict()
    ev["Refine"] = 0
    ev["Degree"] = 1
    if type(geom).__name__ == "TriaMesh":
        ev["Dimension"] = 2
    elif type(geom).__name__ == "TetMesh":
        ev["Dimension"] = 3
    ev["Elements"] = len(geom.t)
    ev["DoF"] = len(geom.v)
    ev["NumEW"] = k
    ev["Eigenvalues"] = evals
    ev["Eigenvectors"] = evecs
    return ev


def compute_shapedna_batch(geoms, k: int = 50, lump: bool = False, workers: int = 0, autotune: bool = False,
                           tune_cache=None, **eig_opts) -> list:
    """ShapeDNA of a collection of meshes in one call: what a loop of
    ``compute_shapedna(geom, k, lump)`` over ``geoms`` returns (lapy/
    shapedna.py:96-166), with the meshes assembled on the GPU and their
    eigenproblems solved concurrently (lh_eigs_batched; ``workers = 0`` uses
    the autotuned or default count, ``autotune=True`` tunes the worker count
    and the SpMM workgroup sizes on this batch first).  ``tune_cache``: a
    JSON file of tuned shapes (lapy_amd.tuning), loaded first and tuned and
    updated only for shapes it lacks.  Anisotropic assembly is not batched:
    use compute_shapedna."""
    if not geoms:
        return []
    ctx = _lib.default_context()
    As, Bs, keep = [], [], []
    for g in geoms:
        name = type(g).__name__
        if name not in ("TriaMesh", "TetMesh"):
            raise ValueError('Geometry type "' + name + '" unknown')
        dm = device_mesh(g)
        kind = _lib.LH_TRIA if name == "TriaMesh" else _lib.LH_TET
        a, b = dm.assemble(kind, lump, np.float64)
        As.append(a)
        Bs.append(b)
        keep.append(dm)
    ks = [k] * len(geoms)
    if tune_cache is not None:
        from . import tuning

        tuning.autotune_cached(As, Bs, ks, tune_cache, ctx, **eig_opts)
    elif autotune:
        _lib.eigs_autotune(ctx, As, Bs, ks, **eig_opts)
    res = _lib.eigs_batched(ctx, As, Bs, ks, workers=workers, **eig_opts)
    return [_shapedna_dict(g, k, r[0], r[1]) for g, r in zip(geoms, res)]
