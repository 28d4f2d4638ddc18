/*
 * lapyhip.h -- C ABI of the MI355X (gfx950) FEM hot-path library.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes.  No
 * C++ exception crosses this boundary.  Loaded from Python with ctypes by
 * lapy_amd/_lib.py.
 *
 * Each entry point names the reference interface it replaces (LaPy, a pure
 * Python package; the arithmetic it delegates to numpy/scipy is what these
 * functions compute on the GPU).
 *
 * Threading: one lh_ctx per host thread; a ctx owns one HIP stream and is
 * not re-entrant.  Handles created from a ctx belong to it.
 */
#ifndef LAPYHIP_H
#define LAPYHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define LH_OK              0
#define LH_ERR_ARG         1   /* invalid argument (Python: ValueError)      */
#define LH_ERR_OOM         2   /* device allocation failed                   */
#define LH_ERR_HIP         3   /* HIP runtime error                          */
#define LH_ERR_NOCONV      4   /* iterative solver did not converge          */
#define LH_ERR_NUMERIC     5   /* breakdown (non-SPD Gram, NaN, ...)         */

/* ---- dtypes -------------------------------------------------------------- */
#define LH_F64 0
#define LH_F32 1

/* ---- element kinds ------------------------------------------------------- */
#define LH_TRIA        0   /* lapy/solver.py:106-194  Solver._fem_tria          */
#define LH_TRIA_ANISO  1   /* lapy/solver.py:197-308  Solver._fem_tria_aniso    */
#define LH_TRIA_MASS   2   /* lapy/solver.py:311-377  Solver.fem_tria_mass      */
#define LH_TET         3   /* lapy/solver.py:380-533  Solver._fem_tetra         */

typedef struct lh_ctx  lh_ctx;   /* device, stream, scratch, last error        */
typedef struct lh_mesh lh_mesh;  /* device-resident mesh + cached topology     */
typedef struct lh_csr  lh_csr;   /* device-resident CSR matrix (n x n)         */
typedef struct lh_eig  lh_eig;   /* device-resident eigen result               */
typedef struct lh_mpnn lh_mpnn;  /* learned preconditioner (weights on device) */

/* Per-solve statistics returned by the solvers. */
typedef struct lh_stats {
    int32_t iterations;      /* outer iterations                              */
    int32_t converged;       /* number of converged wanted pairs / rhs        */
    int32_t fallback_used;   /* 1 if the learned preconditioner was replaced  */
    int32_t precision_mode;  /* LH_PREC_*                                     */
    double  max_resid;       /* largest final relative residual               */
    double  t_setup_us;      /* host-measured phase times (microseconds)      */
    double  t_iter_us;
    double  t_rr_us;         /* eigensolver: host time in the Rayleigh-Ritz
                                calls; those inside the iteration are issued
                                without waiting (their results are read with
                                the next residual norms)                      */
    double  t_total_us;
    int32_t spmm_calls;
    int32_t rr_calls;
    int32_t restarts;
    int32_t iterations_fp32; /* LH_PREC_MIXED: iterations of the FP32 phase
                                (included in `iterations`)                    */
    int32_t poly_columns;    /* eigensolver: block columns the Chebyshev
                                polynomial ran on, summed over iterations
                                (converged wanted pairs get plain Jacobi)     */
} lh_stats;

#define LH_PREC_FP64   0   /* FP64 everywhere                                */
#define LH_PREC_MIXED  1   /* an FP32 phase (FP32 blocks and SpMM, FP32 matrix-
                              core Grams and updates, FP64 Rayleigh-Ritz and
                              Cholesky factors) until its residuals stop
                              falling or reach 1e-5, then FP64 iterations from
                              its X block to the requested tolerance        */

#define LH_PRECOND_JACOBI  0   /* Jacobi with the automatic Chebyshev polynomial */
#define LH_PRECOND_LEARNED 1   /* the context's network (lh_ctx_set_mpnn)        */
#define LH_PRECOND_NONE    2
#define LH_PRECOND_AUTO    3   /* eigensolver: the network where the automatic
                                  polynomial is plain Jacobi (small meshes) and
                                  a network is set, else LH_PRECOND_JACOBI      */

/* storage of the Chebyshev preconditioner's work blocks (arithmetic is FP32) */
#define LH_STORE_AUTO 0
#define LH_STORE_F32  1
#define LH_STORE_F16  2   /* columns scaled to a fixed RMS, saturating stores */

/* ---- context ------------------------------------------------------------- */
/* New; the reference runs on the host and has no device context. */
int  lh_ctx_create(int device, lh_ctx** out);
void lh_ctx_destroy(lh_ctx* ctx);
const char* lh_last_error(lh_ctx* ctx);
int  lh_sync(lh_ctx* ctx);
/* Number of visible HIP devices (0 when no GPU is present). */
int  lh_device_count(void);

/* ---- mesh ---------------------------------------------------------------- */
/* Uploads the arrays LaPy's Solver reads from a mesh: geometry.v (nv x 3,
 * float32 or float64 -- the element math runs in this dtype, as numpy does at
 * lapy/solver.py:145-165) and geometry.t (ne x nodes, int64), see
 * lapy/tria_mesh.py:238-276 and lapy/tet_mesh.py:51-59.  Indices are checked
 * on the device against [0, nv) before any kernel gathers with them. */
int lh_mesh_create(lh_ctx* ctx, const void* v, int v_dtype, int64_t nv,
                   const int64_t* t, int64_t ne, int nodes, lh_mesh** out);
/* Same, but v / t are already device pointers (v: x,y,z SoA planes of nv
 * values each, t: int32 ne x nodes row-major); they are copied. */
int lh_mesh_create_device(lh_ctx* ctx, const void* v_soa, int v_dtype, int64_t nv,
                          const int32_t* t_dev, int64_t ne, int nodes, lh_mesh** out);
/* Device-to-device copy of a mesh's v and t without its cached topology (the
 * next assembly of the clone runs the symbolic phase again). */
int lh_mesh_clone(lh_mesh* src, lh_mesh** out);
void lh_mesh_free(lh_mesh* m);
/* Matrix dimension n = max(t)+1 (scipy's inferred csc shape,
 * lapy/solver.py:178) and pattern size; builds the topology on first use. */
int lh_mesh_pattern_info(lh_mesh* m, int64_t* n, int64_t* nnz);
/* Average edge length: lapy/tria_mesh.py:735-748, lapy/tet_mesh.py:182-195. */
int lh_mesh_avg_edge_length(lh_mesh* m, double* out);

/* ---- assembly -------------------------------------------------------------
 * Replaces the numpy element math and scipy's COO->CSC duplicate summation of
 * lapy/solver.py:106-194 (LH_TRIA), :197-308 (LH_TRIA_ANISO), :311-377
 * (LH_TRIA_MASS, mass only, A_out ignored) and :380-533 (LH_TET).
 * out_dtype is the Solver's dtype (lapy/solver.py:61).  lump selects the
 * diagonal mass (lapy/solver.py:188-193).  For LH_TRIA_ANISO, u1/u2
 * (ne x 3) and aniso_mat (ne x 2) are host arrays in v's dtype
 * (lapy/solver.py:76-91); else pass NULL.
 * The pattern (indptr/indices) equals scipy's exactly; values are summed in
 * ascending element order, so A is bit-symmetric. */
int lh_assemble(lh_ctx* ctx, lh_mesh* m, int kind, int lump, int out_dtype,
                const void* u1, const void* u2, const void* aniso_mat,
                lh_csr** A_out, lh_csr** B_out);
/* Numeric phase only (topology reused), results written into existing CSR
 * handles from an earlier lh_assemble of the same mesh and kind. */
int lh_assemble_into(lh_ctx* ctx, lh_mesh* m, int kind, int lump,
                     const void* u1, const void* u2, const void* aniso_mat,
                     lh_csr* A, lh_csr* B);
/* Row order of the triangle assembly (LH_TRIA, LH_TRIA_MASS; no counterpart
 * in the reference, results are identical in every order): the rows are
 * processed in the mesh's vertex order, or in the pattern's locality order
 * (the vertices' Morton codes sorted, as the solvers use) with the coordinates gathered from a copy in that order
 * (each row still written at its own CSR offset and summed in ascending
 * element order).  LH_ASM_ORDER_AUTO takes the locality order for meshes of
 * at least 65,536 used vertices (fewer than 2^24, rows of at most 64
 * entries) from their second numeric assembly on, or once a solve has built
 * the pattern's order (building the order costs more than one assembly);
 * LH_ASM_ORDER_LOCALITY takes it whenever the mesh is eligible. */
#define LH_ASM_ORDER_AUTO     0
#define LH_ASM_ORDER_MESH     1
#define LH_ASM_ORDER_LOCALITY 2
int lh_ctx_set_assembly_order(lh_ctx* ctx, int order);

/* ---- gradient / divergence on triangles and tetrahedra ----------------------
 * mode LH_DG_GRADIENT: in = nf vertex functions (nv x nf, float64, row-major),
 *   out = per-element gradients (ne x nf x 3) -- lapy/diffgeo.py:222-300
 *   (tria_compute_gradient) and :846-918 (tet_compute_gradient); nf = 1 is
 *   the (n,) -> (ne, 3) form.
 * mode LH_DG_DIVERGENCE: in = per-element fields (ne x nf x 3), out = their
 *   integrated divergence at the vertices (nv x nf) -- lapy/diffgeo.py:303-387
 *   and :921-1006.
 * mode LH_DG_GEODESIC_RHS: in = nf vertex functions, out = the integrated
 *   divergence of their normalised gradients (nan_to_num'd), the right-hand
 *   side that compute_geodesic_f hands to Solver.poisson, lapy/diffgeo.py:117-164.
 * Geometry arithmetic runs in v's dtype, the function arithmetic in float64. */
enum { LH_DG_GRADIENT = 0, LH_DG_DIVERGENCE = 1, LH_DG_GEODESIC_RHS = 2 };
int lh_diffgeo(lh_ctx* ctx, lh_mesh* m, int mode, const double* in, int nf, double* out);

/* ---- principal curvatures (triangle meshes) ---------------------------------
 * lh_curvature replaces TriaMesh.curvature (lapy/tria_mesh.py:1078-1219):
 * per vertex (nv rows, host outputs, any may be NULL) the unit principal
 * directions u_min, u_max (nv x 3 row-major), the curvatures c_min <= c_max,
 * c_mean = (c_min + c_max) / 2 and c_gauss = c_min c_max (taken before the
 * two are ordered, as in the reference), and the normals (nv x 3), after
 * `smoothit` neighbour-averaging steps of the vertex tensors.
 * lh_curvature_tria replaces TriaMesh.curvature_tria (lapy/tria_mesh.py:
 * 1221-1279): per triangle u_min, u_max (ne x 3) and c_min, c_max (ne), the
 * input of LH_TRIA_ANISO.  Eigenvector signs: the largest-magnitude
 * component is positive; a triangle's corners' u_min are turned to agree
 * with the first corner's before they are averaged (the reference keeps
 * LAPACK's signs).  FP64 arithmetic; the mesh must be oriented. */
int lh_curvature(lh_ctx* ctx, lh_mesh* m, int smoothit, double* u_min, double* u_max, double* c_min,
                 double* c_max, double* c_mean, double* c_gauss, double* normals);
int lh_curvature_tria(lh_ctx* ctx, lh_mesh* m, int smoothit, double* u_min, double* u_max, double* c_min,
                      double* c_max);

/* ---- CSR matrices ----------------------------------------------------------- */
/* A CSR of a symmetric matrix equals its CSC, which is how the façade hands
 * back scipy.sparse.csc_matrix (lapy/solver.py:100-101). */
int  lh_csr_info(lh_csr* a, int64_t* n, int64_t* nnz, int* dtype, int* is_diag);
int  lh_csr_download(lh_csr* a, int32_t* indptr, int32_t* indices, void* data);
/* Upload a host CSR (e.g. a mass matrix a caller replaced with sparse.eye,
 * lapy/diffgeo.py:149).  data dtype LH_F64 or LH_F32. */
int  lh_csr_create(lh_ctx* ctx, int64_t n, int64_t nnz, const int32_t* indptr,
                   const int32_t* indices, const void* data, int dtype, lh_csr** out);
void lh_csr_free(lh_csr* a);
/* y = a * x for nrhs row-major columns (x, y host arrays, n x nrhs, f64). */
int  lh_csr_matmul(lh_ctx* ctx, lh_csr* a, const double* x, int nrhs, double* y);

/* ---- generalized eigensolve  A x = lambda B x -------------------------------- */
typedef struct lh_eig_opts {
    int32_t precision;     /* LH_PREC_*                                        */
    int32_t precond;       /* LH_PRECOND_*                                     */
    int32_t block;         /* block size m >= k (0: automatic)                 */
    int32_t max_iter;      /* 0: automatic                                     */
    double  tol;           /* relative residual target (0: automatic)          */
    uint64_t seed;         /* start block seed                                 */
    int32_t verbose;
    int32_t poly_degree;   /* with LH_PRECOND_JACOBI: degree of the Chebyshev
                              polynomial in D^-1 (A - sigma B) applied on top
                              of Jacobi; 0: automatic (~sqrt(n)/6.3, at most
                              128), 1: plain Jacobi                          */
    double  poly_ratio;    /* Chebyshev interval [lmax/ratio, lmax]
                              (0: automatic, 2 * degree^2)                   */
    int32_t poly_storage;  /* LH_STORE_* for the Chebyshev work blocks        */
    int32_t no_fallback;   /* LH_PRECOND_LEARNED: 1 keeps the network even when
                              the geometric mean of the wanted pairs' absolute
                              residuals has not halved over the last 8
                              iterations (0: switch to plain Jacobi then,
                              stats.fallback_used = 1)                        */
} lh_eig_opts;

/* Replaces Solver.eigs / eigensystem (lapy/solver.py:667-716: splu of
 * A - sigma*B, then eigsh shift-invert).  Returns the k eigenpairs nearest
 * sigma for sigma below the spectrum (the smallest k), ascending, vectors
 * B-orthonormal.  Results stay on the device in *out. */
int lh_eigs(lh_ctx* ctx, lh_csr* A, lh_csr* B, int k, double sigma,
            const lh_eig_opts* opts, lh_eig** out, lh_stats* stats);
/* Many independent problems A_i x = lambda B_i x, k_i eigenpairs each (the
 * reference computes a shape collection's ShapeDNA one mesh at a time,
 * lapy/shapedna.py:96-166).  The problems are taken largest first by
 * `workers` concurrent solvers, each with its own stream and a device arena
 * for its buffers (so no allocation synchronises the device between
 * problems); workers = 0 uses the count lh_eigs_autotune chose for this
 * context (8 when it was not run).  out[i] receives problem i's result
 * (free with lh_eig_free); status[i] (nullable) its status; stats[i]
 * (nullable) its statistics.  Returns LH_OK, or the first failure's code
 * (LH_ERR_NOCONV: some problem stopped at max_iter; its result is still in
 * out[i]). */
int lh_eigs_batched(lh_ctx* ctx, int nmesh, lh_csr* const* A, lh_csr* const* B, const int* k, double sigma,
                    const lh_eig_opts* opts, int workers, lh_eig** out, int* status, lh_stats* stats);
/* Autotuner of the batched solver (no counterpart in the reference): solves
 * the given problems with 1, 2, 4, 8 and 16 workers, keeps the fastest count
 * in the context (used by lh_eigs_batched with workers = 0) and reports it
 * and the five wall times (seconds, nullable); then, for each problem shape
 * not yet in the context's table, times the ELL SpMM pair on the solver's
 * block width with 64, 128 and 256 threads per workgroup and keeps the
 * fastest (used by every solve on the context and its batch workers). */
int lh_eigs_autotune(lh_ctx* ctx, int nmesh, lh_csr* const* A, lh_csr* const* B, const int* k, double sigma,
                     const lh_eig_opts* opts, int* best_workers, double* seconds);
/* The context's tuned launch shapes, for a cache kept across processes
 * (lapy_amd.tuning stores them keyed by lh_device_arch): one entry per
 * matrix shape (n, nnz), with the k and block width it was tuned on and the
 * SpMM workgroup size (64, 128 or 256).  lh_tune_set adds an entry or
 * replaces the one of the same shape; none of these may run while a solve on
 * the context does.  lh_ctx_get_workers / lh_ctx_set_workers: the tuned
 * worker count of lh_eigs_batched (0: not tuned).  lh_device_arch: the
 * device's architecture name (e.g. "gfx950:sramecc+:xnack-"). */
typedef struct {
    int64_t n;
    int64_t nnz;
    int32_t k;
    int32_t ncol;
    int32_t block;
    int32_t reserved;
} lh_tune_entry;
int lh_tune_count(lh_ctx* ctx, int* count);
int lh_tune_get(lh_ctx* ctx, int i, lh_tune_entry* out);
int lh_tune_set(lh_ctx* ctx, const lh_tune_entry* e);
int lh_ctx_get_workers(lh_ctx* ctx, int* workers);
int lh_ctx_set_workers(lh_ctx* ctx, int workers);
int lh_device_arch(lh_ctx* ctx, char* buf, int len);
/* Copy results to host: evals (k), evecs (n x k, column-major = numpy
 * Fortran order, i.e. evecs[:, j] is contiguous). Either may be NULL. */
int  lh_eig_download(lh_eig* e, double* evals, double* evecs);
void lh_eig_free(lh_eig* e);

/* ---- linear solves ((alpha*B + beta*A) x = b) ----------------------------------
 * Replaces the splu solves of heat.diffusion (lapy/heat.py:204-227, alpha=1,
 * beta=t, B lumped) and Solver.poisson (lapy/solver.py:843-877, alpha=0,
 * beta=1).  dmask (n bytes, nullable): 1 marks a Dirichlet row/column that is
 * removed from the system (lapy/solver.py:848-865); those entries of x are
 * left as given in x on input.  b and x are host, n x nrhs column-major.
 * With no Dirichlet rows and alpha == 0 the operator is singular (Neumann):
 * the right-hand side and the iterates are projected onto the complement of
 * the constant vector and the mean-free solution is returned. */
typedef struct lh_pcg_opts {
    int32_t precond;       /* LH_PRECOND_*                                      */
    int32_t max_iter;      /* 0: automatic                                      */
    double  rtol;          /* ||r|| / ||b||                                     */
    int32_t verbose;
    int32_t no_fallback;   /* LH_PRECOND_LEARNED: 1 keeps the network; 0 falls
                              back to Jacobi (restarting CG) when the residual
                              falls by less than half per iteration over three
                              4-iteration convergence checks running         */
    int32_t poly_degree;   /* LH_PRECOND_JACOBI: degree of the Chebyshev-Jacobi
                              polynomial z = p(D^-1 K) D^-1 r (symmetric
                              positive definite, on [bmax / (2 deg^2), bmax],
                              bmax the Gershgorin bound of D^-1 K); 0 or 1:
                              plain Jacobi (the default: on one right-hand
                              side each Chebyshev step costs ~0.4 of a
                              Jacobi-CG iteration, so fewer iterations save
                              little time); at most 256                      */
    int32_t reserved;
} lh_pcg_opts;
int lh_pcg(lh_ctx* ctx, lh_csr* A, lh_csr* B, double alpha, double beta,
           const double* b, int nrhs, const uint8_t* dmask, double* x,
           const lh_pcg_opts* opts, lh_stats* stats);

/* ---- measurement helpers (no counterpart in the reference) -------------------
 * Average duration in microseconds of one launch of a hot kernel, timed with
 * HIP events on the context's stream over `iters` back-to-back launches.
 * lh_time_assembly: the numeric assembly kernel of lh_assemble_into (topology
 * already built, A/B from an lh_assemble of the same mesh/kind/lump).
 * lh_time_spmm: the fused A/B SpMM on an n x ncol row-major block (FP64 or
 * FP32 per `precision`). */
int lh_time_assembly(lh_ctx* ctx, lh_mesh* m, int kind, int lump, lh_csr* A, lh_csr* B, int iters,
                     double* avg_us);
int lh_time_spmm(lh_ctx* ctx, lh_csr* A, lh_csr* B, int ncol, int precision, int iters, double* avg_us);
/* One step of the eigensolver's Chebyshev-Jacobi preconditioner (FP32
 * arithmetic, fused SpMM + update, work blocks stored per `storage` =
 * LH_STORE_F32 or LH_STORE_F16) on an n x ncol block, optionally after the
 * solver's locality reordering (reorder = 1, as lh_eigs does). */
int lh_time_cheb_step(lh_ctx* ctx, lh_csr* A, lh_csr* B, int ncol, int iters, int reorder, int storage,
                      double* avg_us);
/* ---- learned preconditioner (no counterpart in the reference) -------------
 * A 3-layer message-passing network of width 32 on the operator's graph,
 * linear in the residual: for K = beta A + alpha B, s = diag(K)^-1/2, the
 * normalised off-diagonal weights W^ = S (K - diag K) S and per-vertex
 * features phi_i = (1, log(d_i / mean d), log(mu_i / mean mu),
 * sum_j W^_ij) (d = diag K, mu = row sums of B; 0 without B), gates
 * gam_l = 1 + tanh(phi Gam_l):
 *   H0 = (s r) u^T,  H_{l+1} = gam_l * (H_l Ws_l^T + (W^ H_l) Wn_l^T),
 *   Q x = H3 v,  z = s (s r + Q^T Q s r)
 * (Jacobi plus a learned correction, symmetric positive definite; v = 0 is
 * exactly Jacobi).  Weights and hidden states are bfloat16 on the device,
 * the layer products run on the matrix cores (FP32 accumulation).
 * Parameters (lh_mpnn_param_count() floats): u (32), then Ws_l, Wn_l
 * (32 x 32 row-major each) for l = 0, 1, 2, then v (32), then Gam_l (4 x 32
 * row-major) for l = 0, 1, 2.  params NULL: seeded small random weights with
 * Gam = 0 (gates 1).
 * lh_ctx_set_mpnn selects the network LH_PRECOND_LEARNED uses in lh_eigs
 * (K = A - sigma B) and lh_pcg (K = alpha B + beta A); NULL clears it.
 * lh_mpnn_apply: Z = P R for host R, Z (n x nc row-major).
 * lh_mpnn_loss_grad: training loss L = (1/nc) sum_c ||S(K z_c - r_c)||^2 /
 * ||S r_c||^2 and dL/dparams (host, FP64).
 * lh_mpnn_time_layer: average duration (us) of one forward layer on an
 * n x nc block (HIP events on the context's stream, `iters` launches), in
 * the eigensolver's locality order when reorder = 1 (as the solver runs it). */
int  lh_mpnn_param_count(void);
int  lh_mpnn_create(lh_ctx* ctx, const float* params, lh_mpnn** out);
void lh_mpnn_free(lh_mpnn* net);
int  lh_mpnn_set_params(lh_mpnn* net, const float* params);
int  lh_mpnn_get_params(lh_mpnn* net, float* params);
int  lh_ctx_set_mpnn(lh_ctx* ctx, lh_mpnn* net);
int  lh_mpnn_apply(lh_mpnn* net, lh_csr* A, lh_csr* B, double alpha, double beta, const double* R, int nc,
                   double* Z);
int  lh_mpnn_loss_grad(lh_mpnn* net, lh_csr* A, lh_csr* B, double alpha, double beta, const double* R, int nc,
                       double* loss, double* grad);
int  lh_mpnn_time_layer(lh_mpnn* net, lh_csr* A, lh_csr* B, double alpha, double beta, int nc, int iters,
                        int reorder, double* avg_us);

/* The eigensolver's tall-skinny block products on their own (no counterpart
 * in the reference, whose eigsh does its dense work inside ARPACK; exposed
 * for their tests and timing), on host row-major buffers:
 * lh_block_gram: C (p x q, row-major) = U^T V over n rows; flags:
 *   LH_GRAM_SYM: the product is known to be symmetric, p == q; only upper
 *   tiles are formed;
 * lh_block_gemm: Y = alpha U C + beta Y with C p x q row-major; flags:
 *   LH_GEMM_UPPER: C is upper triangular (its zero blocks are skipped),
 *   LH_GEMM_INPLACE: Y <- alpha Y C in place (U unused, p == q <= 256,
 *   beta == 0), the eigensolver's Cholesky-QR and Ritz rotations;
 * both: LH_BLOCK_F32: the blocks are rounded to FP32 on the device and the
 *   FP32 matrix-core kernels run (the mixed-precision eigensolver's), C of
 *   the Gram summed in FP64, Y returned widened to FP64.
 * iters > 0 and avg_us non-NULL: also the average kernel time of `iters`
 * further runs (lh_block_gemm times them with beta = 0, after Y is returned). */
#define LH_GRAM_SYM     1
#define LH_GEMM_UPPER   1
#define LH_GEMM_INPLACE 2
#define LH_BLOCK_F32    4
int lh_block_gram(lh_ctx* ctx, int64_t n, const double* U, int ldu, int p, const double* V, int ldv, int q,
                  int flags, double* C, int iters, double* avg_us);
int lh_block_gemm(lh_ctx* ctx, int64_t n, int p, int q, const double* U, int ldu, const double* Cs, double* Y,
                  int ldy, double alpha, double beta, int flags, int iters, double* avg_us);
/* lh_block_gram without any allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  U (n x p,
 * leading dimension ldu), V (n x q, ldv) and C (p x q, row-major, contiguous)
 * are DEVICE pointers to FP64 blocks; flags: 0 or LH_GRAM_SYM.  The product
 * is launched on the context's stream and the call returns without waiting
 * (lh_sync, or any later synchronising call, completes it).  The per-wave
 * partial tiles live in one fixed 16 MiB device array per GPU
 * (512 tiles of 64 x 64 doubles), so p and q are limited to
 * LH_GRAM_NOALLOC_MAX_COLS (larger: LH_ERR_ARG); n is not limited (more rows
 * per wave).  All calls on one GPU share that array: calls that overlap on
 * one GPU (several contexts' streams or threads at once) race.  The rows are
 * split over fewer waves than in lh_block_gram, so the two agree to rounding,
 * not bit for bit. */
#define LH_GRAM_NOALLOC_MAX_COLS 1024
int lh_block_gram_noalloc(lh_ctx* ctx, int64_t n, const double* U, int ldu, int p, const double* V, int ldv, int q,
                          int flags, double* C);
/* The Rayleigh-Ritz eigensolver on its own (no counterpart in the reference;
 * exposed for its tests and timing): smallest nev eigenpairs of the symmetric
 * N x N host matrix H (N <= 192).  evals (nev), Z (N x nev row-major); when
 * phase_us is not NULL it receives 7 phase durations in microseconds. */
int lh_small_symeig(lh_ctx* ctx, int N, int nev, const double* H, double* evals, double* Z, double* phase_us);
/* The eigensolver's Cholesky-QR factor on its own (no counterpart in the
 * reference; exposed for its tests): the upper triangular m x m T with
 * T^T G T = I for the symmetric positive definite m x m host matrix G
 * (m <= 128, row-major, shift added to the equilibrated diagonal), computed
 * on one workgroup as in the block solver.  min_pivot (nullable) receives the
 * smallest pivot of the equilibrated matrix.  LH_ERR_NUMERIC when G is not
 * numerically positive definite (a pivot <= 1e-12 relative to its unit
 * diagonal). */
int lh_small_cholinv(lh_ctx* ctx, int m, const double* G, double shift, double* T, double* min_pivot);

#ifdef __cplusplus
}
#endif
#endif /* LAPYHIP_H */
