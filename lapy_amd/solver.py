"""Drop-in ``Solver``: LaPy's FEM solver API on the MI355X.

Mirrors lapy/solver.py:14-889: ``Solver(geometry, lump, aniso, aniso_smooth,
use_cholmod, dtype)`` assembles ``stiffness`` / ``mass`` on the GPU
(lh_assemble), ``eigs`` / ``eigensystem`` solve ``A x = lambda B x`` on the GPU
(lh_eigs, block LOBPCG) and ``poisson`` solves with PCG (lh_pcg).  The
matrices stay on the device; ``.stiffness`` / ``.mass`` hand back
``scipy.sparse.csc_matrix`` copies on first access and may be reassigned
(e.g. ``fem.mass = sparse.eye(n)``, lapy/diffgeo.py:149), in which case the
new matrix is uploaded when a solve needs it.
"""

from __future__ import annotations

import logging

import numpy as np
from scipy import sparse

from . import _lib
from .mesh import device_mesh

logger = logging.getLogger(__name__)


def import_optional_dependency(name: str, extra: str = "", raise_error: bool = True):
    """lapy/utils/_imports.py:15-63 (same ImportError text)."""
    import importlib

    install_name = {"sksparse": "scikit-sparse"}.get(name, name)
    try:
        return importlib.import_module(name)
    except ImportError:
        if raise_error:
            raise ImportError(
                f"Missing optional dependency '{install_name}'. {extra} "
                f"Use pip or conda to install {install_name}."
            ) from None
        return None


class _MatrixSlot:
    """A matrix that lives on the device, with a lazily downloaded host copy."""

    def __init__(self, dev=None, host=None):
        self.dev = dev
        self.host = host

    def get_host(self):
        if self.host is None:
            self.host = self.dev.to_scipy_csc()
        return self.host

    def get_dev(self, ctx, dtype=np.float64):
        if self.dev is None or self.dev.dtype != dtype:
            self.dev = _lib.DeviceCSR.from_scipy(ctx, self.get_host(), dtype=dtype)
        return self.dev


class Solver:
    """FEM solver for the Laplace eigenvalue and Poisson equations (GPU).

    Parameters are those of lapy/solver.py:54-62.  The eigensolver
    options (``precision``, ``precond``, ``tol``, ``block``) are keyword-only
    extras with defaults that keep the reference's behaviour (FP64).
    """

    def __init__(self, geometry, lump: bool = False, aniso=None, aniso_smooth: int = 10,
                 use_cholmod: bool = False, dtype=np.float64) -> None:
        if use_cholmod:
            self.sksparse = import_optional_dependency("sksparse", raise_error=True)
        else:
            self.sksparse = None
        dtype = np.dtype(dtype)
        if dtype not in (np.dtype(np.float64), np.dtype(np.float32)):
            raise ValueError(f"dtype must be float64 or float32, got {dtype}")
        name = type(geometry).__name__
        if name not in ("TriaMesh", "TetMesh"):
            raise ValueError('Geometry type "' + name + '" unknown')
        self._ctx = _lib.default_context()
        if name == "TriaMesh":
            dm = device_mesh(geometry, self._ctx)
            if aniso is not None:
                logger.info("TriaMesh with anisotropic Laplace-Beltrami")
                u1, u2, c1, c2 = geometry.curvature_tria(smoothit=aniso_smooth)
                if isinstance(aniso, (list, tuple, set, np.ndarray)):
                    if len(aniso) != 2:
                        raise ValueError("aniso should be scalar or tuple/array of length 2!")
                    aniso0, aniso1 = list(aniso)
                else:
                    aniso0 = aniso1 = aniso
                aniso_mat = np.empty((geometry.t.shape[0], 2), dtype=dtype)
                aniso_mat[:, 1] = np.exp(-aniso1 * np.abs(c1))
                aniso_mat[:, 0] = np.exp(-aniso0 * np.abs(c2))
                a, b = dm.assemble(_lib.LH_TRIA_ANISO, lump, dtype, u1, u2, aniso_mat)
            else:
                logger.info("TriaMesh with regular Laplace-Beltrami")
                a, b = dm.assemble(_lib.LH_TRIA, lump, dtype)
        elif name == "TetMesh":
            logger.info("TetMesh with regular Laplace")
            dm = device_mesh(geometry, self._ctx)
            a, b = dm.assemble(_lib.LH_TET, lump, dtype)
        else:
            raise ValueError('Geometry type "' + name + '" unknown')
        self._a = _MatrixSlot(dev=a)
        self._b = _MatrixSlot(dev=b)
        self.geotype = type(geometry)
        self.use_cholmod = use_cholmod
        self.last_stats = None

    # -- matrices ----------------------------------------------------------
    @property
    def stiffness(self):
        return self._a.get_host()

    @stiffness.setter
    def stiffness(self, value):
        self._a = _MatrixSlot(host=value)

    @property
    def mass(self):
        return self._b.get_host()

    @mass.setter
    def mass(self, value):
        self._b = _MatrixSlot(host=value)

    def device_matrices(self, dtype=np.float64):
        """(A, B) as device CSR handles (uploads reassigned matrices)."""
        return self._a.get_dev(self._ctx, dtype), self._b.get_dev(self._ctx, dtype)

    # -- static mass-only assembly --------------------------------------------
    @staticmethod
    def fem_tria_mass(tria, lump: bool = False, dtype=np.float64):
        """lapy/solver.py:311-377 on the GPU -> scipy.sparse.csc_matrix."""
        dm = device_mesh(tria)
        _, b = dm.assemble(_lib.LH_TRIA_MASS, lump, dtype)
        return b.to_scipy_csc()

    # -- eigensolve -----------------------------------------------------------
    def eigs(self, k: int = 10, sigma: float = -0.01, *, precision: str = "fp64", precond: str = "jacobi",
             tol: float = 0.0, block: int = 0, max_iter: int = 0, seed: int = 0, poly_degree: int = 0,
             poly_storage: str = "auto"):
        """k eigenpairs of ``A x = lambda B x`` nearest ``sigma`` (lapy/solver.py:667-716).

        Returns ``(eigenvalues (k,), eigenvectors (n, k))``, ascending, vectors
        B-orthonormal.  For ``sigma`` below the spectrum (the reference default
        -0.01 on a Laplacian) these are the k smallest eigenpairs; for a
        ``sigma`` inside the spectrum the smallest pairs are computed until the
        k nearest ``sigma`` are among them.
        """
        if self.use_cholmod:
            logger.info("Solver: Cholesky decomposition from scikit-sparse cholmod ...")
        else:
            logger.info("Solver: GPU block LOBPCG ...")
        a, b = self.device_matrices(np.float64)
        n = a.n
        k = int(k)
        if k <= 0 or k >= n:
            raise ValueError(f"k must be between 1 and ndim(A)-1 (k={k}, n={n})")
        prec = {"fp64": _lib.LH_PREC_FP64, "mixed": _lib.LH_PREC_MIXED}[precision]
        pc = {"jacobi": _lib.LH_PRECOND_JACOBI, "learned": _lib.LH_PRECOND_LEARNED,
              "none": _lib.LH_PRECOND_NONE, "auto": _lib.LH_PRECOND_AUTO}[precond]
        store = {"auto": _lib.LH_STORE_AUTO, "f32": _lib.LH_STORE_F32, "f16": _lib.LH_STORE_F16}[poly_storage]
        kk = k
        while True:
            evals, evecs, st = _lib.eigs(self._ctx, a, b, kk, sigma, precision=prec, precond=pc, block=block,
                                         max_iter=max_iter, tol=tol, seed=seed, poly_degree=poly_degree,
                                         poly_storage=store)
            self.last_stats = st
            if sigma <= evals[0] or kk >= n - 1:
                break
            # sigma inside the spectrum: the k nearest must be bracketed
            dist = np.abs(evals - sigma)
            order = np.argsort(dist, kind="stable")[:k]
            if evals[-1] - sigma >= dist[order].max() or kk >= n - 1:
                break
            kk = min(n - 1, 2 * kk)
        if kk != k:
            dist = np.abs(evals - sigma)
            sel = np.sort(np.argsort(dist, kind="stable")[:k])
            evals, evecs = evals[sel], evecs[:, sel]
        return evals, evecs

    eigensystem = eigs

    # -- Poisson ----------------------------------------------------------------
    def poisson(self, h=0.0, dtup=(), ntup=()):
        """Solve ``A x = B h`` with Dirichlet / Neumann data (lapy/solver.py:718-889).

        Same argument checks and return shapes as the reference; the linear
        solve is PCG on the GPU (Dirichlet rows removed by masking).
        """
        stiff_slot, mass_slot = self._a, self._b
        a_dev = stiff_slot.get_dev(self._ctx, np.float64)
        m_dev = mass_slot.get_dev(self._ctx, np.float64)
        dim = a_dev.n
        dtype = np.dtype(self._dtype_of_stiffness())
        if m_dev.n != dim:
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
        didx = np.array([], dtype=np.int64)
        dvals = None
        if dtup:
            if len(dtup) != 2:
                raise ValueError("dtup should contain index and data arrays")
            didx = np.asarray(dtup[0])
            ddat = np.asarray(dtup[1])
            if np.unique(didx).size != len(didx):
                raise ValueError("dtup indices need to be unique")
            if not (len(didx) > 0 and len(didx) == len(ddat)):
                raise ValueError("dtup should contain index and data arrays (same lengths > 0)")
            dvals = np.zeros((dim, n_rhs), dtype=np.float64)
            dvals[didx, :] = np.asarray(ddat, dtype=dtype)[:, None]
        nvec = 0.0
        if ntup:
            if len(ntup) != 2:
                raise ValueError("ntup should contain index and data arrays")
            nidx, ndat = np.asarray(ntup[0]), np.asarray(ntup[1])
            if not (len(nidx) > 0 and len(nidx) == len(ndat)):
                raise ValueError("ntup should contain index and data arrays (same lengths > 0)")
            nvec = np.zeros((dim, n_rhs), dtype=np.float64)
            np.add.at(nvec, (np.tile(nidx, n_rhs), np.repeat(np.arange(n_rhs), len(nidx))),
                      np.tile(np.asarray(ndat, dtype=np.float64), n_rhs))
        # right-hand side b = B (h - n) - A d, products on the GPU
        rhs = m_dev.matmul(np.asarray(h, dtype=np.float64) - nvec)
        if dvals is not None:
            rhs = rhs - a_dev.matmul(dvals)
        dmask = None
        x0 = None
        if len(didx) > 0:
            dmask = np.zeros(dim, dtype=np.uint8)
            dmask[didx] = 1
            x0 = dvals
        x, st = _lib.pcg(self._ctx, a_dev, None, 0.0, 1.0, rhs, x0=x0, dmask=dmask, rtol=1e-12)
        self.last_stats = st
        if len(didx) > 0:
            x[didx, :] = dvals[didx, :]
        x = x.astype(dtype, copy=False)
        if scalar_rhs:
            return np.squeeze(np.array(x))
        return np.array(x)

    def _dtype_of_stiffness(self):
        if self._a.host is not None:
            return self._a.host.dtype
        return self._a.dev.dtype
