// Rayleigh-Ritz for wide trial bases (N > 192 columns, k > 53 wanted pairs):
// the generalized problem GA c = theta GM c is handed to rocSOLVER's
// divide-and-conquer sygvd on the context's stream.  The LDS solvers of
// lh_rr.hip cover the narrower bases of the usual k <= 53; the one-workgroup
// global-memory solver (k_rr) they fall back to costs ~0.4 s at N = 720.
#include "lh_dense.hpp"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

namespace lh {

namespace {

// A = (G + G^T) / 2 (row- or column-major alike) for both Grams
__global__ void k_sym_pair(int N, const double* __restrict__ GA, const double* __restrict__ GM,
                           double* __restrict__ A, double* __restrict__ B) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)N * N) return;
    const int i = (int)(idx / N), j = (int)(idx % N);
    A[idx] = 0.5 * (GA[(int64_t)i * N + j] + GA[(int64_t)j * N + i]);
    B[idx] = 0.5 * (GM[(int64_t)i * N + j] + GM[(int64_t)j * N + i]);
}

// C (N x nev, row-major) = the first nev columns of V (N x N, column-major);
// theta = the lowest nev eigenvalues (D ascending)
__global__ void k_take_lowest(int N, int nev, const double* __restrict__ V, const double* __restrict__ D,
                              double* __restrict__ C, double* __restrict__ theta) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)N * nev) return;
    const int i = (int)(idx / nev), j = (int)(idx % nev);
    C[idx] = V[(int64_t)j * N + i];
    if (i == 0) theta[j] = D[j];
}

void check_blas(rocblas_status s, const char* what) {
    if (s != rocblas_status_success) LH_THROW(LH_ERR_HIP, "%s failed (rocblas status %d)", what, (int)s);
}

}  // namespace

size_t rr_wide_work_doubles(int N) { return 2 * (size_t)N * N + 2 * (size_t)N; }

void rr_solve_wide(lh_ctx* ctx, int N, int nev, const double* GA, const double* GM, double* C, double* theta,
                   double* work, int* status) {
    if (N < 1 || nev < 1 || nev > N) LH_THROW(LH_ERR_ARG, "wide Rayleigh-Ritz: N = %d, nev = %d", N, nev);
    if (!ctx->blas) {
        rocblas_handle h = nullptr;
        check_blas(rocblas_create_handle(&h), "rocblas_create_handle");
        ctx->blas = h;
    }
    rocblas_handle h = static_cast<rocblas_handle>(ctx->blas);
    check_blas(rocblas_set_stream(h, ctx->stream), "rocblas_set_stream");
    double* A = work;
    double* B = A + (size_t)N * N;
    double* D = B + (size_t)N * N;
    double* E = D + N;
    k_sym_pair<<<nblocks((int64_t)N * N, 256), 256, 0, ctx->stream>>>(N, GA, GM, A, B);
    LH_CHECK_LAUNCH();
    // status[0] = info: 0 on success, > 0 when GM is not positive definite or
    // the tridiagonal solver did not converge
    check_blas(rocsolver_dsygvd(h, rocblas_eform_ax, rocblas_evect_original, rocblas_fill_upper, N, A, N, B, N, D, E,
                                reinterpret_cast<rocblas_int*>(status)),
               "rocsolver_dsygvd");
    k_take_lowest<<<nblocks((int64_t)N * nev, 256), 256, 0, ctx->stream>>>(N, nev, A, D, C, theta);
    LH_CHECK_LAUNCH();
}

}  // namespace lh

void release_blas(lh_ctx* ctx) {
    if (ctx && ctx->blas) {
        (void)rocblas_destroy_handle(static_cast<rocblas_handle>(ctx->blas));
        ctx->blas = nullptr;
    }
}
