"""Drop-in ``heat.diffusion`` (lapy/heat.py:114-232) on the GPU.

Backward Euler ``(M_lumped + t A) u = b0`` with ``t = m * avg_edge_length^2``
and ``b0`` the 0/1 indicator of each seed set; solved with PCG on the device
(all seed sets at once) instead of SuperLU.
"""

from __future__ import annotations

import logging

import numpy as np

from . import _lib
from .solver import Solver, import_optional_dependency

logger = logging.getLogger(__name__)


def diffusion(geometry, vids, m: float = 1.0, aniso=None, use_cholmod: bool = False, *, rtol: float = 1e-12,
              precond: str = "jacobi"):
    """Heat diffusion from the vertices in ``vids`` (same arguments, errors and
    return shapes as lapy/heat.py:114-232).  Keyword-only additions: ``rtol``
    of the PCG solve and ``precond`` ("jacobi", or "learned" for the network
    installed on the context, which falls back to Jacobi when CG stalls)."""
    pc = {"jacobi": _lib.LH_PRECOND_JACOBI, "learned": _lib.LH_PRECOND_LEARNED}[precond]
    if use_cholmod:
        import_optional_dependency("sksparse", raise_error=True)
    nv = len(geometry.v)
    if isinstance(vids, list) and len(vids) > 0 and isinstance(vids[0], (list, np.ndarray)):
        scalar_input = False
        vids_list = [np.asarray(v, dtype=int).ravel() for v in vids]
    else:
        scalar_input = True
        vids_list = [np.asarray(vids, dtype=int).ravel()]
    for v in vids_list:
        if np.any(v < 0) or np.any(v >= nv):
            raise ValueError("vids contains out-of-range vertex indices")
    fem = Solver(geometry, lump=True, aniso=aniso)
    t = m * geometry.avg_edge_length() ** 2
    a, b = fem.device_matrices(np.float64)
    n = a.n
    b0 = np.zeros((n, len(vids_list)))
    for k, v in enumerate(vids_list):
        b0[v, k] = 1.0
    logger.info("Solver: PCG on the GPU")
    vfunc, st = _lib.pcg(fem._ctx, a, b, 1.0, t, b0, rtol=rtol, precond=pc)
    fem.last_stats = st
    if n < nv:
        # trailing vertices outside every element: the matrix is smaller
        out = np.zeros((nv, len(vids_list)))
        out[:n] = vfunc
        vfunc = out
    if scalar_input:
        return vfunc[:, 0]
    return vfunc
