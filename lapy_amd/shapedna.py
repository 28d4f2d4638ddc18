"""Drop-in ``shapedna.compute_shapedna`` (lapy/shapedna.py:96-166)."""

from __future__ import annotations

import numpy as np

from .solver import Solver


def compute_shapedna(geom, k: int = 50, lump: bool = False, aniso=None, aniso_smooth: int = 10,
                     use_cholmod: bool = False, **eig_opts) -> dict:
    """Eigenvalues / eigenvectors of the Laplace-Beltrami operator as a dict
    with the reference's keys."""
    fem = Solver(geom, lump=lump, aniso=aniso, aniso_smooth=aniso_smooth, use_cholmod=use_cholmod)
    evals, evecs = fem.eigs(k=k, **eig_opts)
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
