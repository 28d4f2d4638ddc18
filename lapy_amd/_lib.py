"""ctypes binding of liblapyhip.so (the C ABI declared in include/lapyhip.h).

The library is required: importing the package works without it (so that
``build()`` can run), but every call that needs the GPU raises
``RuntimeError`` when the shared library is missing or no HIP device is
visible.  There is no CPU fallback.
"""

from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "liblapyhip.so")

LH_OK, LH_ERR_ARG, LH_ERR_OOM, LH_ERR_HIP, LH_ERR_NOCONV, LH_ERR_NUMERIC = range(6)
LH_F64, LH_F32 = 0, 1
LH_TRIA, LH_TRIA_ANISO, LH_TRIA_MASS, LH_TET = 0, 1, 2, 3
LH_ASM_ORDER_AUTO, LH_ASM_ORDER_MESH, LH_ASM_ORDER_LOCALITY = 0, 1, 2
LH_DG_GRADIENT, LH_DG_DIVERGENCE, LH_DG_GEODESIC_RHS = 0, 1, 2
LH_PREC_FP64, LH_PREC_MIXED = 0, 1
LH_PRECOND_JACOBI, LH_PRECOND_LEARNED, LH_PRECOND_NONE, LH_PRECOND_AUTO = 0, 1, 2, 3
LH_STORE_AUTO, LH_STORE_F32, LH_STORE_F16 = 0, 1, 2

_vp = C.c_void_p
_i32 = C.c_int32
_i64 = C.c_int64
_dp = C.POINTER(C.c_double)


class Stats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32),
        ("converged", C.c_int32),
        ("fallback_used", C.c_int32),
        ("precision_mode", C.c_int32),
        ("max_resid", C.c_double),
        ("t_setup_us", C.c_double),
        ("t_iter_us", C.c_double),
        ("t_rr_us", C.c_double),
        ("t_total_us", C.c_double),
        ("spmm_calls", C.c_int32),
        ("rr_calls", C.c_int32),
        ("restarts", C.c_int32),
        ("iterations_fp32", C.c_int32),
        ("poly_columns", C.c_int32),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class EigOpts(C.Structure):
    _fields_ = [
        ("precision", C.c_int32),
        ("precond", C.c_int32),
        ("block", C.c_int32),
        ("max_iter", C.c_int32),
        ("tol", C.c_double),
        ("seed", C.c_uint64),
        ("verbose", C.c_int32),
        ("poly_degree", C.c_int32),
        ("poly_ratio", C.c_double),
        ("poly_storage", C.c_int32),
        ("no_fallback", C.c_int32),
    ]


class PcgOpts(C.Structure):
    _fields_ = [
        ("precond", C.c_int32),
        ("max_iter", C.c_int32),
        ("rtol", C.c_double),
        ("verbose", C.c_int32),
        ("no_fallback", C.c_int32),
        ("poly_degree", C.c_int32),
        ("reserved", C.c_int32),
    ]


# name -> (restype, argtypes); every symbol of include/lapyhip.h
SIGNATURES = {
    "lh_device_count": (C.c_int, []),
    "lh_ctx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "lh_ctx_destroy": (None, [_vp]),
    "lh_last_error": (C.c_char_p, [_vp]),
    "lh_sync": (C.c_int, [_vp]),
    "lh_mesh_create": (C.c_int, [_vp, _vp, C.c_int, _i64, _vp, _i64, C.c_int, C.POINTER(_vp)]),
    "lh_mesh_create_device": (C.c_int, [_vp, _vp, C.c_int, _i64, _vp, _i64, C.c_int, C.POINTER(_vp)]),
    "lh_mesh_clone": (C.c_int, [_vp, C.POINTER(_vp)]),
    "lh_mesh_free": (None, [_vp]),
    "lh_mesh_pattern_info": (C.c_int, [_vp, C.POINTER(_i64), C.POINTER(_i64)]),
    "lh_mesh_avg_edge_length": (C.c_int, [_vp, _dp]),
    "lh_diffgeo": (C.c_int, [_vp, _vp, C.c_int, _vp, C.c_int, _vp]),
    "lh_assemble": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, C.POINTER(_vp), C.POINTER(_vp)]),
    "lh_assemble_into": (C.c_int, [_vp, _vp, C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp]),
    "lh_csr_info": (C.c_int, [_vp, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "lh_csr_download": (C.c_int, [_vp, _vp, _vp, _vp]),
    "lh_csr_create": (C.c_int, [_vp, _i64, _i64, _vp, _vp, _vp, C.c_int, C.POINTER(_vp)]),
    "lh_csr_free": (None, [_vp]),
    "lh_csr_matmul": (C.c_int, [_vp, _vp, _vp, C.c_int, _vp]),
    "lh_eigs": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_double, C.POINTER(EigOpts), C.POINTER(_vp), C.POINTER(Stats)]),
    "lh_eig_download": (C.c_int, [_vp, _vp, _vp]),
    "lh_eig_free": (None, [_vp]),
    "lh_pcg": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_double, _vp, C.c_int, _vp, _vp, C.POINTER(PcgOpts),
                         C.POINTER(Stats)]),
    "lh_time_assembly": (C.c_int, [_vp, _vp, C.c_int, C.c_int, _vp, _vp, C.c_int, _dp]),
    "lh_time_spmm": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _dp]),
    "lh_small_symeig": (C.c_int, [_vp, C.c_int, C.c_int, _vp, _vp, _vp, _vp]),
    "lh_small_cholinv": (C.c_int, [_vp, C.c_int, _vp, C.c_double, _vp, _dp]),
    "lh_time_cheb_step": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, _dp]),
    "lh_eigs_batched": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, C.c_double, _vp, C.c_int, _vp, _vp, _vp]),
    "lh_eigs_autotune": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, C.c_double, _vp, _vp, _vp]),
    "lh_tune_count": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "lh_tune_get": (C.c_int, [_vp, C.c_int, _vp]),
    "lh_tune_set": (C.c_int, [_vp, _vp]),
    "lh_ctx_get_workers": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "lh_ctx_set_workers": (C.c_int, [_vp, C.c_int]),
    "lh_device_arch": (C.c_int, [_vp, C.c_char_p, C.c_int]),
    "lh_curvature": (C.c_int, [_vp, _vp, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "lh_curvature_tria": (C.c_int, [_vp, _vp, C.c_int, _vp, _vp, _vp, _vp]),
    "lh_mpnn_param_count": (C.c_int, []),
    "lh_mpnn_create": (C.c_int, [_vp, _vp, _vp]),
    "lh_mpnn_free": (None, [_vp]),
    "lh_mpnn_set_params": (C.c_int, [_vp, _vp]),
    "lh_mpnn_get_params": (C.c_int, [_vp, _vp]),
    "lh_ctx_set_mpnn": (C.c_int, [_vp, _vp]),
    "lh_ctx_set_assembly_order": (C.c_int, [_vp, C.c_int]),
    "lh_mpnn_apply": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_double, _vp, C.c_int, _vp]),
    "lh_mpnn_loss_grad": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_double, _vp, C.c_int, _dp, _vp]),
    "lh_mpnn_time_layer": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, _dp]),
    "lh_block_gram": (C.c_int, [_vp, C.c_int64, _vp, C.c_int, C.c_int, _vp, C.c_int, C.c_int, C.c_int, _vp, C.c_int,
                                _dp]),
    "lh_block_gram_noalloc": (C.c_int, [_vp, C.c_int64, _vp, C.c_int, C.c_int, _vp, C.c_int, C.c_int, C.c_int, _vp]),
    "lh_block_gemm": (C.c_int, [_vp, C.c_int64, C.c_int, C.c_int, _vp, C.c_int, _vp, _vp, C.c_int, C.c_double,
                                C.c_double, C.c_int, C.c_int, _dp]),
}

_lib = None
_lock = threading.RLock()  # default_context() -> Context() -> lib() re-enters


def lib():
    """Load liblapyhip.so (raises RuntimeError when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"liblapyhip.so not found at {LIB_PATH}; build it with "
                    "`python -c 'import __graft_entry__ as g; g.build()'` (hipcc, gfx950)")
            h = C.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                f = getattr(h, name)
                f.restype = res
                f.argtypes = args
            _lib = h
    return _lib


class LapyHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class NoConvergence(LapyHipError):
    """Raised when an iterative solve stops before reaching its tolerance."""


def check(code, ctx=None, what=""):
    if code == LH_OK:
        return
    msg = lib().lh_last_error(ctx.handle if ctx is not None else None)
    msg = (msg.decode() if msg else "") or f"error {code}"
    if what:
        msg = f"{what}: {msg}"
    if code == LH_ERR_ARG:
        raise ValueError(msg)
    if code == LH_ERR_NOCONV:
        raise NoConvergence(code, msg)
    raise LapyHipError(code, msg)


class TuneEntry(C.Structure):
    """lh_tune_entry: a tuned ELL SpMM workgroup size for one matrix shape."""
    _fields_ = [("n", C.c_int64), ("nnz", C.c_int64), ("k", C.c_int32), ("ncol", C.c_int32), ("block", C.c_int32),
                ("reserved", C.c_int32)]


class Context:
    """One HIP device + stream (lh_ctx)."""

    def __init__(self, device: int = 0):
        L = lib()
        nd = L.lh_device_count()
        if nd <= 0:
            raise RuntimeError("no HIP device visible: the lapy_amd GPU path needs an MI355X (gfx950)")
        h = _vp()
        check(L.lh_ctx_create(device, C.byref(h)), None, "lh_ctx_create")
        self.handle = h
        self.device = device

    def __del__(self):
        try:
            if self.handle:
                lib().lh_ctx_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    def sync(self):
        check(lib().lh_sync(self.handle), self, "lh_sync")

    def arch(self) -> str:
        """The device's architecture name (lh_device_arch)."""
        buf = C.create_string_buffer(256)
        check(lib().lh_device_arch(self.handle, buf, 256), self, "lh_device_arch")
        return buf.value.decode()

    def tuning(self):
        """(tuned worker count or 0, [dict(n, nnz, k, ncol, block), ...])."""
        w = C.c_int(0)
        check(lib().lh_ctx_get_workers(self.handle, C.byref(w)), self, "lh_ctx_get_workers")
        cnt = C.c_int(0)
        check(lib().lh_tune_count(self.handle, C.byref(cnt)), self, "lh_tune_count")
        out = []
        for i in range(cnt.value):
            e = TuneEntry()
            check(lib().lh_tune_get(self.handle, i, C.byref(e)), self, "lh_tune_get")
            out.append({"n": e.n, "nnz": e.nnz, "k": e.k, "ncol": e.ncol, "block": e.block})
        return w.value, out

    def set_tuning(self, workers: int, entries):
        """Install a tuned worker count and SpMM shapes (as returned by tuning())."""
        check(lib().lh_ctx_set_workers(self.handle, int(workers)), self, "lh_ctx_set_workers")
        for d in entries:
            e = TuneEntry(int(d["n"]), int(d["nnz"]), int(d["k"]), int(d["ncol"]), int(d["block"]), 0)
            check(lib().lh_tune_set(self.handle, C.byref(e)), self, "lh_tune_set")

    def set_assembly_order(self, order: int):
        """Row order of the triangle assembly (LH_ASM_ORDER_*); the values
        do not depend on it."""
        check(lib().lh_ctx_set_assembly_order(self.handle, int(order)), self, "lh_ctx_set_assembly_order")


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        with _lock:
            if _default_ctx is None:
                _default_ctx = Context(int(os.environ.get("LAPY_AMD_DEVICE", "0")))
    return _default_ctx


def _ptr(a):
    return a.ctypes.data_as(_vp) if a is not None else None


class DeviceCSR:
    """Device-resident CSR matrix (lh_csr); symmetric, so its CSR is its CSC."""

    def __init__(self, ctx: Context, handle):
        self.ctx = ctx
        self.handle = handle
        n, nnz, dt, isd = _i64(), _i64(), C.c_int(), C.c_int()
        check(lib().lh_csr_info(handle, C.byref(n), C.byref(nnz), C.byref(dt), C.byref(isd)), ctx)
        self.n, self.nnz = n.value, nnz.value
        self.dtype = np.float64 if dt.value == LH_F64 else np.float32
        self.is_diag = bool(isd.value)

    @classmethod
    def from_scipy(cls, ctx: Context, mat, dtype=None):
        from scipy import sparse

        m = sparse.csr_matrix(mat)
        m.sum_duplicates()
        m.sort_indices()
        if dtype is None:
            dtype = np.float32 if m.dtype == np.float32 else np.float64
        data = np.ascontiguousarray(m.data, dtype=dtype)
        indptr = np.ascontiguousarray(m.indptr, dtype=np.int32)
        indices = np.ascontiguousarray(m.indices, dtype=np.int32)
        if m.shape[0] != m.shape[1]:
            raise ValueError("matrix must be square")
        h = _vp()
        check(lib().lh_csr_create(ctx.handle, m.shape[0], indices.size, _ptr(indptr), _ptr(indices), _ptr(data),
                                  LH_F64 if dtype == np.float64 else LH_F32, C.byref(h)), ctx, "lh_csr_create")
        return cls(ctx, h)

    def download(self):
        indptr = np.empty(self.n + 1, dtype=np.int32)
        indices = np.empty(self.nnz, dtype=np.int32)
        data = np.empty(self.nnz, dtype=self.dtype)
        check(lib().lh_csr_download(self.handle, _ptr(indptr), _ptr(indices), _ptr(data)), self.ctx)
        return indptr, indices, data

    def to_scipy_csc(self):
        """scipy.sparse.csc_matrix of this (symmetric) matrix."""
        from scipy import sparse

        indptr, indices, data = self.download()
        m = sparse.csc_matrix((data, indices, indptr), shape=(self.n, self.n))
        m.has_sorted_indices = True
        return m

    def matmul(self, x):
        x = np.asarray(x, dtype=np.float64)
        squeeze = x.ndim == 1
        x2 = x[:, None] if squeeze else x
        xr = np.ascontiguousarray(x2)  # row-major n x nrhs
        y = np.empty_like(xr)
        check(lib().lh_csr_matmul(self.ctx.handle, self.handle, _ptr(xr), xr.shape[1], _ptr(y)), self.ctx)
        return y[:, 0] if squeeze else y

    def __del__(self):
        try:
            if self.handle:
                lib().lh_csr_free(self.handle)
                self.handle = None
        except Exception:
            pass


class DeviceMesh:
    """Device-resident mesh (lh_mesh) with its cached topology."""

    def __init__(self, ctx: Context, v, t):
        v = np.asarray(v)
        t = np.asarray(t)
        if v.dtype == np.float32:
            vd, vdt = np.ascontiguousarray(v, dtype=np.float32), LH_F32
        else:
            vd, vdt = np.ascontiguousarray(v, dtype=np.float64), LH_F64
        if vd.ndim != 2 or vd.shape[1] != 3:
            raise ValueError("vertices must have shape (n, 3)")
        td = np.ascontiguousarray(t, dtype=np.int64)
        if td.ndim != 2 or td.shape[1] not in (3, 4):
            raise ValueError("elements must have shape (m, 3) or (m, 4)")
        h = _vp()
        check(lib().lh_mesh_create(ctx.handle, _ptr(vd), vdt, vd.shape[0], _ptr(td), td.shape[0], td.shape[1],
                                   C.byref(h)), ctx, "lh_mesh_create")
        self.ctx = ctx
        self.handle = h
        self.nodes = td.shape[1]
        self.vdtype = vd.dtype
        self.nv, self.ne = vd.shape[0], td.shape[0]

    def clone(self):
        """Device-to-device copy without the cached topology."""
        h = _vp()
        check(lib().lh_mesh_clone(self.handle, C.byref(h)), self.ctx, "lh_mesh_clone")
        other = DeviceMesh.__new__(DeviceMesh)
        other.ctx, other.handle, other.nodes, other.vdtype = self.ctx, h, self.nodes, self.vdtype
        other.nv, other.ne = self.nv, self.ne
        return other

    def time_assembly(self, kind, lump, A, B, iters=20):
        us = C.c_double()
        check(lib().lh_time_assembly(self.ctx.handle, self.handle, kind, int(bool(lump)),
                                     A.handle if A is not None else None, B.handle, iters, C.byref(us)), self.ctx)
        return us.value

    def pattern_info(self):
        n, nnz = _i64(), _i64()
        check(lib().lh_mesh_pattern_info(self.handle, C.byref(n), C.byref(nnz)), self.ctx)
        return n.value, nnz.value

    def avg_edge_length(self):
        out = C.c_double()
        check(lib().lh_mesh_avg_edge_length(self.handle, C.byref(out)), self.ctx)
        return out.value

    def diffgeo(self, mode, data, nf):
        """Gradient (mode 0), divergence (1) or geodesic right-hand side (2) on
        a triangle mesh; data is (nv, nf) for modes 0/2, (ne, nf, 3) for 1."""
        d = np.ascontiguousarray(data, dtype=np.float64)
        ne, nv = self.ne, self.nv
        out = np.empty((ne, nf, 3) if mode == LH_DG_GRADIENT else (nv, nf), dtype=np.float64)
        check(lib().lh_diffgeo(self.ctx.handle, self.handle, int(mode), _ptr(d), int(nf), _ptr(out)),
              self.ctx, "lh_diffgeo")
        return out

    def assemble(self, kind, lump, dtype=np.float64, u1=None, u2=None, aniso_mat=None):
        a, b = _vp(), _vp()
        odt = LH_F64 if np.dtype(dtype) == np.float64 else LH_F32
        keep = []

        def arr(x, cols):
            if x is None:
                return None
            y = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, cols)
            keep.append(y)
            return _ptr(y)

        pa = arr(u1, 3)
        pb = arr(u2, 3)
        pc = arr(aniso_mat, 2)
        check(lib().lh_assemble(self.ctx.handle, self.handle, kind, int(bool(lump)), odt, pa, pb, pc,
                                C.byref(a) if kind != LH_TRIA_MASS else None, C.byref(b)), self.ctx, "lh_assemble")
        A = DeviceCSR(self.ctx, a) if kind != LH_TRIA_MASS else None
        return A, DeviceCSR(self.ctx, b)

    def __del__(self):
        try:
            if self.handle:
                lib().lh_mesh_free(self.handle)
                self.handle = None
        except Exception:
            pass


def eigs(ctx: Context, A: DeviceCSR, B: DeviceCSR, k: int, sigma: float = -0.01, precision=LH_PREC_FP64,
         precond=LH_PRECOND_JACOBI, block=0, max_iter=0, tol=0.0, seed=0, verbose=0, keep_device=False,
         poly_degree=0, poly_ratio=0.0, poly_storage=LH_STORE_AUTO, no_fallback=False, allow_noconv=False):
    """Generalized eigenpairs on the GPU; returns (evals, evecs, stats).
    allow_noconv: return the Ritz pairs of the last iteration when max_iter
    stops the solve (stats.converged < k) instead of raising NoConvergence."""
    opts = EigOpts(precision, precond, block, max_iter, tol, seed, verbose, poly_degree, poly_ratio, poly_storage,
                   int(bool(no_fallback)))
    st = Stats()
    h = _vp()
    code = lib().lh_eigs(ctx.handle, A.handle, B.handle, int(k), float(sigma), C.byref(opts), C.byref(h), C.byref(st))
    try:
        if not (allow_noconv and code == LH_ERR_NOCONV and h):
            check(code, ctx, "lh_eigs")
        evals = np.empty(k, dtype=np.float64)
        if keep_device:
            check(lib().lh_eig_download(h, _ptr(evals), None), ctx)
            return evals, EigHandle(ctx, h), st
        evecs_t = np.empty((k, A.n), dtype=np.float64)
        check(lib().lh_eig_download(h, _ptr(evals), _ptr(evecs_t)), ctx)
        return evals, evecs_t.T, st
    finally:
        if h and not keep_device:
            lib().lh_eig_free(h)


def _eig_opts(precision=LH_PREC_FP64, precond=LH_PRECOND_JACOBI, block=0, max_iter=0, tol=0.0, seed=0, verbose=0,
              poly_degree=0, poly_ratio=0.0, poly_storage=LH_STORE_AUTO, no_fallback=False):
    return EigOpts(precision, precond, block, max_iter, tol, seed, verbose, poly_degree, poly_ratio, poly_storage,
                   int(bool(no_fallback)))


def _problem_arrays(As, Bs, ks):
    if not (len(As) == len(Bs) == len(ks)):
        raise ValueError("As, Bs and ks must have the same length")
    nm = len(As)
    pa = (C.c_void_p * max(nm, 1))(*[a.handle for a in As])
    pb = (C.c_void_p * max(nm, 1))(*[b.handle for b in Bs])
    pk = (C.c_int * max(nm, 1))(*[int(x) for x in ks])
    return nm, pa, pb, pk


def eigs_batched(ctx: Context, As, Bs, ks, sigma: float = -0.01, workers: int = 0, keep_device: bool = False,
                 **opts):
    """Many independent generalized eigenproblems on the GPU, solved by
    concurrent solver streams (lh_eigs_batched).  Returns a list of
    (evals, evecs | EigHandle, stats, status) in the input order; raises on
    argument errors, and NoConvergence when some problem stopped at max_iter
    (after all results were gathered, with the list as its .results)."""
    nm, pa, pb, pk = _problem_arrays(As, Bs, ks)
    o = _eig_opts(**opts)
    out = (C.c_void_p * max(nm, 1))()
    status = (C.c_int * max(nm, 1))()
    stats = (Stats * max(nm, 1))()
    code = lib().lh_eigs_batched(ctx.handle, nm, pa, pb, pk, float(sigma), C.byref(o), int(workers), out, status,
                                 stats)
    results = []
    for i in range(nm):
        h = out[i]
        if not h:
            results.append(None)
            continue
        k = int(ks[i])
        if keep_device:
            evals = np.empty(k, dtype=np.float64)
            check(lib().lh_eig_download(h, _ptr(evals), None), ctx)
            results.append((evals, EigHandle(ctx, h), stats[i], status[i]))
        else:
            evals = np.empty(k, dtype=np.float64)
            evecs_t = np.empty((k, As[i].n), dtype=np.float64)
            try:
                check(lib().lh_eig_download(h, _ptr(evals), _ptr(evecs_t)), ctx)
            finally:
                lib().lh_eig_free(h)
            results.append((evals, evecs_t.T, stats[i], status[i]))
    if code == LH_ERR_NOCONV:
        err = NoConvergence(code, lib().lh_last_error(ctx.handle).decode())
        err.results = results
        raise err
    check(code, ctx, "lh_eigs_batched")
    return results


def eigs_autotune(ctx: Context, As, Bs, ks, sigma: float = -0.01, **opts):
    """Time the batched solver with 1, 2, 4, 8 and 16 workers on these
    problems; the fastest count is kept in the context.  Returns
    (best_workers, {workers: seconds})."""
    nm, pa, pb, pk = _problem_arrays(As, Bs, ks)
    o = _eig_opts(**opts)
    best = C.c_int(0)
    secs = (C.c_double * 5)()
    check(lib().lh_eigs_autotune(ctx.handle, nm, pa, pb, pk, float(sigma), C.byref(o), C.byref(best), secs), ctx,
          "lh_eigs_autotune")
    return best.value, dict(zip((1, 2, 4, 8, 16), list(secs)))


class Mpnn:
    """Device handle of the learned preconditioner's weights (lh_mpnn)."""

    def __init__(self, ctx: Context, params=None):
        self.ctx = ctx
        self.handle = None
        h = _vp()
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float32)
        if p is not None and p.size != mpnn_param_count():
            raise ValueError(f"expected {mpnn_param_count()} parameters, got {p.size}")
        check(lib().lh_mpnn_create(ctx.handle, _ptr(p), C.byref(h)), ctx, "lh_mpnn_create")
        self.handle = h

    @property
    def params(self):
        out = np.empty(mpnn_param_count(), dtype=np.float32)
        check(lib().lh_mpnn_get_params(self.handle, _ptr(out)), self.ctx)
        return out

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != mpnn_param_count():
            raise ValueError(f"expected {mpnn_param_count()} parameters, got {p.size}")
        check(lib().lh_mpnn_set_params(self.handle, _ptr(p)), self.ctx)

    def apply(self, A: DeviceCSR, B, alpha: float, beta: float, R):
        R = np.ascontiguousarray(np.asarray(R, dtype=np.float64).reshape(A.n, -1))
        Z = np.empty_like(R)
        check(lib().lh_mpnn_apply(self.handle, A.handle, B.handle if B is not None else None, float(alpha),
                                  float(beta), _ptr(R), R.shape[1], _ptr(Z)), self.ctx, "lh_mpnn_apply")
        return Z

    def loss_grad(self, A: DeviceCSR, B, alpha: float, beta: float, R):
        R = np.ascontiguousarray(np.asarray(R, dtype=np.float64).reshape(A.n, -1))
        loss = C.c_double()
        g = np.empty(mpnn_param_count(), dtype=np.float64)
        check(lib().lh_mpnn_loss_grad(self.handle, A.handle, B.handle if B is not None else None, float(alpha),
                                      float(beta), _ptr(R), R.shape[1], C.byref(loss), _ptr(g)), self.ctx,
              "lh_mpnn_loss_grad")
        return loss.value, g

    def time_layer(self, A: DeviceCSR, B, alpha: float, beta: float, ncol: int, iters: int = 20,
                   reorder: bool = True) -> float:
        """Average microseconds of one forward layer on an n x ncol block (in
        the eigensolver's locality order when reorder)."""
        us = C.c_double()
        check(lib().lh_mpnn_time_layer(self.handle, A.handle, B.handle if B is not None else None, float(alpha),
                                       float(beta), int(ncol), int(iters), int(bool(reorder)), C.byref(us)),
              self.ctx, "lh_mpnn_time_layer")
        return us.value

    def use(self, on: bool = True):
        """Make this network the context's LH_PRECOND_LEARNED (or clear it)."""
        check(lib().lh_ctx_set_mpnn(self.ctx.handle, self.handle if on else None), self.ctx)

    def __del__(self):
        try:
            if self.handle:
                lib().lh_mpnn_free(self.handle)
                self.handle = None
        except Exception:
            pass


def mpnn_param_count():
    return int(lib().lh_mpnn_param_count())


def time_cheb_step(ctx: Context, A: DeviceCSR, B: DeviceCSR, ncol: int, iters: int = 50, reorder: bool = True,
                   storage=LH_STORE_F16):
    us = C.c_double()
    check(lib().lh_time_cheb_step(ctx.handle, A.handle, B.handle, ncol, iters, int(reorder), int(storage),
                                  C.byref(us)), ctx)
    return us.value


LH_GRAM_SYM, LH_BLOCK_F32 = 1, 4
LH_GRAM_NOALLOC_MAX_COLS = 1024  # lh_block_gram_noalloc's column limit (its fixed scratch)


def block_gram(ctx: Context, U: np.ndarray, V: np.ndarray, p: int, q: int, sym: bool = False, iters: int = 0,
               f32: bool = False):
    """C = U[:, :p]^T V[:, :q] on the GPU (FP64 matrix cores; f32: the blocks
    rounded to FP32 and the FP32 matrix cores); U, V row-major float64
    (n x ldu), (n x ldv).  Returns (C, average microseconds or None)."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    V = np.ascontiguousarray(V, dtype=np.float64)
    if U.shape[0] != V.shape[0]:
        raise ValueError("U and V must have the same number of rows")
    out = np.empty((p, q), dtype=np.float64)
    us = C.c_double(0.0)
    flags = (LH_GRAM_SYM if sym else 0) | (LH_BLOCK_F32 if f32 else 0)
    check(lib().lh_block_gram(ctx.handle, U.shape[0], _ptr(U), U.shape[1], p, _ptr(V), V.shape[1], q, flags,
                              _ptr(out), iters, C.byref(us) if iters else None), ctx)
    return out, (us.value if iters else None)


LH_GEMM_UPPER, LH_GEMM_INPLACE = 1, 2


def block_gemm(ctx: Context, U, p: int, Cm: np.ndarray, Y: np.ndarray, alpha=1.0, beta=0.0, iters: int = 0,
               upper: bool = False, inplace: bool = False, f32: bool = False):
    """Y[:, :q] = alpha U[:, :p] Cm + beta Y[:, :q] on the GPU (FP64 matrix
    cores; f32: U and Y rounded to FP32, the FP32 matrix cores, Y widened
    back); inplace: Y[:, :p] = alpha Y[:, :p] Cm (U is ignored); upper: Cm is
    upper triangular.  Returns (Y, average microseconds or None)."""
    Cm = np.ascontiguousarray(Cm, dtype=np.float64)
    Y = np.array(Y, dtype=np.float64, order="C", copy=True)
    q = Cm.shape[1]
    if inplace:
        U = None
        if p != q:
            raise ValueError("in-place update needs a square C")
    else:
        U = np.ascontiguousarray(U, dtype=np.float64)
        if U.shape[0] != Y.shape[0]:
            raise ValueError("shape mismatch")
    if Cm.shape[0] != p:
        raise ValueError("shape mismatch")
    flags = (LH_GEMM_UPPER if upper else 0) | (LH_GEMM_INPLACE if inplace else 0) | (LH_BLOCK_F32 if f32 else 0)
    us = C.c_double(0.0)
    check(lib().lh_block_gemm(ctx.handle, Y.shape[0], p, q, _ptr(U), U.shape[1] if U is not None else 0, _ptr(Cm),
                              _ptr(Y), Y.shape[1], float(alpha), float(beta), flags, iters,
                              C.byref(us) if iters else None), ctx)
    return Y, (us.value if iters else None)


def time_spmm(ctx: Context, A: DeviceCSR, B: DeviceCSR, ncol: int, precision=LH_PREC_FP64, iters: int = 20):
    us = C.c_double()
    check(lib().lh_time_spmm(ctx.handle, A.handle, B.handle, ncol, precision, iters, C.byref(us)), ctx)
    return us.value


class EigHandle:
    def __init__(self, ctx, h):
        self.ctx, self.handle = ctx, h

    def download(self, n, k):
        evals = np.empty(k, dtype=np.float64)
        evecs_t = np.empty((k, n), dtype=np.float64)
        check(lib().lh_eig_download(self.handle, _ptr(evals), _ptr(evecs_t)), self.ctx)
        return evals, evecs_t.T

    def __del__(self):
        try:
            if self.handle:
                lib().lh_eig_free(self.handle)
                self.handle = None
        except Exception:
            pass


def pcg(ctx: Context, A: DeviceCSR, B, alpha: float, beta: float, b, x0=None, dmask=None, rtol=1e-12,
        max_iter=0, precond=LH_PRECOND_JACOBI, verbose=0, no_fallback=False, poly_degree=0):
    """Solve (alpha*B + beta*A) x = b (b: (n,) or (n, nrhs)); returns (x, stats).
    poly_degree (Jacobi): Chebyshev-Jacobi polynomial degree (0 or 1: plain Jacobi)."""
    b = np.asarray(b, dtype=np.float64)
    squeeze = b.ndim == 1
    b2 = b[:, None] if squeeze else b
    bf = np.asfortranarray(b2)
    x = np.asfortranarray(np.zeros_like(b2) if x0 is None else np.asarray(x0, dtype=np.float64).reshape(b2.shape))
    dm = None if dmask is None else np.ascontiguousarray(dmask, dtype=np.uint8)
    opts = PcgOpts(precond, max_iter, rtol, verbose, int(bool(no_fallback)), int(poly_degree), 0)
    st = Stats()
    code = lib().lh_pcg(ctx.handle, A.handle, B.handle if B is not None else None, float(alpha), float(beta),
                        _ptr(bf), bf.shape[1], _ptr(dm), _ptr(x), C.byref(opts), C.byref(st))
    check(code, ctx, "lh_pcg")
    return (x[:, 0] if squeeze else np.ascontiguousarray(x)), st


def small_symeig(ctx: Context, H, nev: int, profile: bool = False):
    """Smallest nev eigenpairs of a small symmetric matrix with the GPU
    Rayleigh-Ritz solver (N <= 192); returns (evals, Z[, phase_us])."""
    H = np.ascontiguousarray(H, dtype=np.float64)
    N = H.shape[0]
    ev = np.empty(nev)
    Z = np.empty((N, nev))
    ph = np.zeros(7) if profile else None
    check(lib().lh_small_symeig(ctx.handle, N, nev, _ptr(H), _ptr(ev), _ptr(Z), _ptr(ph)), ctx, "lh_small_symeig")
    return (ev, Z, ph) if profile else (ev, Z)


def small_cholinv(ctx: Context, G, shift: float = 0.0):
    """The block solver's Cholesky-QR factor of a small SPD matrix (m <= 128):
    (T, min_pivot) with T upper triangular and T^T (G + shift D^-2) T = I,
    D = diag(G)^-1/2 (lh_small_cholinv)."""
    G = np.ascontiguousarray(G, dtype=np.float64)
    m = G.shape[0]
    T = np.empty((m, m))
    mp = C.c_double(0.0)
    check(lib().lh_small_cholinv(ctx.handle, m, _ptr(G), float(shift), _ptr(T), C.byref(mp)), ctx,
          "lh_small_cholinv")
    return T, mp.value


def curvature(ctx: Context, dm: "DeviceMesh", smoothit: int):
    """lh_curvature: (u_min, u_max, c_min, c_max, c_mean, c_gauss, normals)."""
    nv = dm.nv
    u_min, u_max, normals = (np.empty((nv, 3)) for _ in range(3))
    c_min, c_max, c_mean, c_gauss = (np.empty(nv) for _ in range(4))
    check(lib().lh_curvature(ctx.handle, dm.handle, int(smoothit), _ptr(u_min), _ptr(u_max), _ptr(c_min), _ptr(c_max),
                             _ptr(c_mean), _ptr(c_gauss), _ptr(normals)), ctx, "lh_curvature")
    return u_min, u_max, c_min, c_max, c_mean, c_gauss, normals


def curvature_tria(ctx: Context, dm: "DeviceMesh", smoothit: int):
    """lh_curvature_tria: (u_min, u_max, c_min, c_max) per triangle."""
    ne = dm.ne
    u_min, u_max = np.empty((ne, 3)), np.empty((ne, 3))
    c_min, c_max = np.empty(ne), np.empty(ne)
    check(lib().lh_curvature_tria(ctx.handle, dm.handle, int(smoothit), _ptr(u_min), _ptr(u_max), _ptr(c_min),
                                  _ptr(c_max)), ctx, "lh_curvature_tria")
    return u_min, u_max, c_min, c_max
