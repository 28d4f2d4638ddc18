// Dense kernels of the block eigensolver.
//
//  * gram:  C = U^T V for tall-skinny U (n x p), V (n x q), row-major blocks.
//    Split over row chunks; each workgroup writes a 64x64 partial, a second
//    kernel adds the partials in fixed chunk order (deterministic).
//  * gemm:  Y = alpha * U C + beta * Y, U tall-skinny, C small.
//  * rr:    one workgroup solves the small dense generalized symmetric
//    eigenproblem GA c = theta GM c of Rayleigh-Ritz in FP64: Cholesky of GM,
//    H = L^-1 GA L^-T, Householder tridiagonalisation, bisection for the
//    wanted (smallest) eigenvalues, inverse iteration for their vectors,
//    back-transformation, C = L^-T Z.
//  * cholinv: T = L^-T for G = L L^T (Cholesky-QR of a block), by symmetric
//    elimination with the Schur complement and T in registers.
//
// These replace the dense steps ARPACK performs inside eigsh
// (lapy/solver.py:713-715); the algorithm is block LOBPCG, not Lanczos.
#include "lh_dense.hpp"

#include <cmath>
#include <type_traits>

namespace lh {

// ---------------------------------------------------------------------------
// Gram C = U^T V and the update Y = alpha U C + beta Y: matrix-core kernels
// of lh_mfma.hip for both FP64 and FP32 blocks
// ---------------------------------------------------------------------------
template <typename T>
void gram(lh_ctx* ctx, int64_t n, const T* U, int ldu, int p, const T* V, int ldv, int q, double* C, int ldc,
          DevBuf& ws, double alpha, bool sym) {
    gram_mfma<T>(ctx, n, U, ldu, p, V, ldv, q, C, ldc, ws, alpha, sym);
}
template void gram<double>(lh_ctx*, int64_t, const double*, int, int, const double*, int, int, double*, int, DevBuf&,
                           double, bool);
template void gram<float>(lh_ctx*, int64_t, const float*, int, int, const float*, int, int, double*, int, DevBuf&,
                          double, bool);

template <typename T>
void gemm_nn(lh_ctx* ctx, int64_t n, int p, int q, const T* U, int ldu, const double* C, int ldc, T* Y, int ldy,
             double alpha, double beta) {
    gemm_mfma<T>(ctx, n, p, q, U, ldu, C, ldc, Y, ldy, alpha, beta);
}
template void gemm_nn<double>(lh_ctx*, int64_t, int, int, const double*, int, const double*, int, double*, int, double,
                              double);
template void gemm_nn<float>(lh_ctx*, int64_t, int, int, const float*, int, const double*, int, float*, int, double,
                             double);

// ---------------------------------------------------------------------------
// single-workgroup dense routines (1024 threads)
// ---------------------------------------------------------------------------
constexpr int RB = 1024;

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
// block-wide sum; red must hold RB/64 doubles; all threads get the result
__device__ double block_sum(double v, double* red) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    const int nw = blockDim.x >> 6;
    for (int i = 0; i < nw; ++i) s += red[i];
    return s;
}

// Cholesky G = L L^T in place (lower, row-major N x N, ld N).  Returns false
// on a non-positive pivot relative to the largest diagonal.
__device__ bool chol_inplace(double* G, int N, double* red, int* flag) {
    const int tid = threadIdx.x;
    if (tid == 0) {
        double mx = 0.0;
        for (int i = 0; i < N; ++i) mx = fmax(mx, fabs(G[i * N + i]));
        red[0] = mx;
    }
    __syncthreads();
    const double dmax = red[0];
    __syncthreads();
    for (int j = 0; j < N; ++j) {
        if (tid == 0) {
            double d = G[j * N + j];
            if (!(d > 1e-14 * dmax)) { flag[0] = 1; d = fmax(fabs(d), 1e-300); }
            G[j * N + j] = sqrt(d);
        }
        __syncthreads();
        const double ljj = G[j * N + j];
        for (int i = j + 1 + tid; i < N; i += RB) G[i * N + j] /= ljj;
        __syncthreads();
        const int L = N - j - 1;
        for (int idx = tid; idx < L * L; idx += RB) {
            int i = j + 1 + idx / L, k = j + 1 + idx % L;
            if (k <= i) G[i * N + k] -= G[i * N + j] * G[k * N + j];
        }
        __syncthreads();
    }
    // zero the strict upper triangle
    for (int idx = tid; idx < N * N; idx += RB) {
        int i = idx / N, k = idx % N;
        if (k > i) G[idx] = 0.0;
    }
    __syncthreads();
    return flag[0] == 0;
}

// Solve L X = B in place on B's columns (N x ncols, ld ldb), L lower (ld N)
__device__ void trsm_lower_cols(const double* L, int N, double* B, int ldb, int ncols) {
    for (int c = threadIdx.x; c < ncols; c += RB) {
        for (int i = 0; i < N; ++i) {
            double s = B[(int64_t)i * ldb + c];
            for (int k = 0; k < i; ++k) s -= L[i * N + k] * B[(int64_t)k * ldb + c];
            B[(int64_t)i * ldb + c] = s / L[i * N + i];
        }
    }
    __syncthreads();
}
// Solve L^T X = B in place on B's columns
__device__ void trsm_lowerT_cols(const double* L, int N, double* B, int ldb, int ncols) {
    for (int c = threadIdx.x; c < ncols; c += RB) {
        for (int i = N - 1; i >= 0; --i) {
            double s = B[(int64_t)i * ldb + c];
            for (int k = i + 1; k < N; ++k) s -= L[k * N + i] * B[(int64_t)k * ldb + c];
            B[(int64_t)i * ldb + c] = s / L[i * N + i];
        }
    }
    __syncthreads();
}

// Householder tridiagonalisation of symmetric H (N x N, ld N) in place.
// On exit d[0..N), e[0..N-1), tau[0..N-2); reflector k is stored in
// H[k+1..N-1][k] (its first entry explicitly).
__device__ void tridiag(double* H, int N, double* d, double* e, double* tau, double* v, double* p, double* red) {
    const int tid = threadIdx.x;
    for (int k = 0; k < N - 2; ++k) {
        const int L = N - k - 1;
        double part = 0.0;
        for (int i = tid; i < L; i += RB) {
            double x = H[(k + 1 + i) * N + k];
            v[i] = x;
            part += x * x;
        }
        const double xn2 = block_sum(part, red);
        const double x0 = v[0];
        double alpha, t;
        if (xn2 == 0.0 || (L == 1)) {
            // nothing to annihilate below the subdiagonal
            if (tid == 0) { d[k] = H[k * N + k]; e[k] = x0; tau[k] = 0.0; }
            __syncthreads();
            continue;
        }
        const double nx = sqrt(xn2);
        alpha = (x0 >= 0.0) ? -nx : nx;
        const double v0 = x0 - alpha;
        const double vn2 = xn2 - x0 * x0 + v0 * v0;
        t = 2.0 / vn2;
        __syncthreads();
        if (tid == 0) {
            v[0] = v0;
            d[k] = H[k * N + k];
            e[k] = alpha;
            tau[k] = t;
            H[(k + 1) * N + k] = v0;
        }
        __syncthreads();
        // p = t * H22 v   (groups of 8 threads per row)
        const int G = 8;
        const int gi = tid / G, gl = tid % G;
        for (int r = gi; r < L; r += RB / G) {
            const double* hr = H + (int64_t)(k + 1 + r) * N + (k + 1);
            double s = 0.0;
            for (int j = gl; j < L; j += G) s += hr[j] * v[j];
            for (int off = G / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
            if (gl == 0) p[r] = t * s;
        }
        __syncthreads();
        double pp = 0.0;
        for (int i = tid; i < L; i += RB) pp += p[i] * v[i];
        const double pv = block_sum(pp, red);
        const double K = 0.5 * t * pv;
        for (int i = tid; i < L; i += RB) p[i] -= K * v[i];  // p := w
        __syncthreads();
        for (int idx = tid; idx < L * L; idx += RB) {
            int i = idx / L, j = idx % L;
            H[(int64_t)(k + 1 + i) * N + (k + 1 + j)] -= v[i] * p[j] + p[i] * v[j];
        }
        __syncthreads();
    }
    if (tid == 0) {
        if (N >= 2) {
            d[N - 2] = H[(N - 2) * N + (N - 2)];
            e[N - 2] = H[(N - 1) * N + (N - 2)];
            tau[N - 2] = 0.0;
        }
        d[N - 1] = H[(N - 1) * N + (N - 1)];
    }
    __syncthreads();
}

// number of eigenvalues of T(d, e2 = e^2) strictly below x
__device__ int sturm_count(const double* d, const double* e2, int N, double x, double pivmin) {
    int cnt = 0;
    double q = d[0] - x;
    if (fabs(q) < pivmin) q = -pivmin;
    if (q < 0) ++cnt;
    for (int i = 1; i < N; ++i) {
        q = (d[i] - x) - e2[i - 1] / q;
        if (fabs(q) < pivmin) q = -pivmin;
        if (q < 0) ++cnt;
    }
    return cnt;
}

// Solve (T - lam I) y = y in place by LU with partial pivoting; ws holds 5N.
__device__ void tri_solve_shifted(const double* d, const double* e, int N, double lam, double tiny, double* y,
                                  double* ws) {
    double* u0 = ws;
    double* u1 = ws + N;
    double* u2 = ws + 2 * N;
    double* lm = ws + 3 * N;
    double* pv = ws + 4 * N;
    double c0 = d[0] - lam, c1 = (N > 1) ? e[0] : 0.0, c2 = 0.0;
    for (int i = 0; i < N - 1; ++i) {
        const double sub = e[i];
        const double a1 = d[i + 1] - lam;
        const double b1 = (i + 1 < N - 1) ? e[i + 1] : 0.0;
        if (fabs(c0) >= fabs(sub)) {
            double l = (c0 != 0.0) ? sub / c0 : 0.0;
            u0[i] = (c0 != 0.0) ? c0 : tiny;
            u1[i] = c1;
            u2[i] = c2;
            lm[i] = l;
            pv[i] = 0.0;
            c0 = a1 - l * c1;
            c1 = b1 - l * c2;
            c2 = 0.0;
        } else {
            double l = c0 / sub;
            u0[i] = sub;
            u1[i] = a1;
            u2[i] = b1;
            lm[i] = l;
            pv[i] = 1.0;
            c0 = c1 - l * a1;
            c1 = c2 - l * b1;
            c2 = 0.0;
        }
    }
    u0[N - 1] = (fabs(c0) > tiny) ? c0 : tiny;
    // forward
    for (int i = 0; i < N - 1; ++i) {
        if (pv[i] != 0.0) { double tmp = y[i]; y[i] = y[i + 1]; y[i + 1] = tmp; }
        y[i + 1] -= lm[i] * y[i];
    }
    // back
    y[N - 1] /= u0[N - 1];
    if (N >= 2) y[N - 2] = (y[N - 2] - u1[N - 2] * y[N - 1]) / u0[N - 2];
    for (int i = N - 3; i >= 0; --i) y[i] = (y[i] - u1[i] * y[i + 1] - u2[i] * y[i + 2]) / u0[i];
}

// Smallest nev eigenpairs of T (d, e): lam[0..nev), Z (N x nev, ld nev).
__device__ void tri_eig_lowest(const double* d, const double* e, int N, int nev, double* lam, double* Z, double* e2,
                               double* ws /* nev * 6N */, double* red) {
    const int tid = threadIdx.x;
    for (int i = tid; i < N - 1; i += RB) e2[i] = e[i] * e[i];
    if (tid == 0) {
        double lo = d[0], hi = d[0], nrm = 0.0;
        for (int i = 0; i < N; ++i) {
            double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < N - 1 ? fabs(e[i]) : 0.0);
            lo = fmin(lo, d[i] - r);
            hi = fmax(hi, d[i] + r);
            nrm = fmax(nrm, fabs(d[i]) + r);
        }
        red[0] = lo;
        red[1] = hi;
        red[2] = fmax(nrm, 1e-300);
    }
    __syncthreads();
    const double glo = red[0], ghi = red[1], tnorm = red[2];
    const double pivmin = 1e-300 + 4.0 * 2.2e-308 * tnorm;
    __syncthreads();
    // bisection, one thread per wanted eigenvalue
    for (int j = tid; j < nev; j += RB) {
        double a = glo - 1e-14 * tnorm, b = ghi + 1e-14 * tnorm;
        for (int it = 0; it < 200; ++it) {
            double mid = 0.5 * (a + b);
            if (mid <= a || mid >= b) break;
            if (b - a <= 2.2e-16 * (fabs(a) + fabs(b)) + 1e-17 * tnorm) break;
            if (sturm_count(d, e2, N, mid, pivmin) > j) b = mid;
            else a = mid;
        }
        lam[j] = 0.5 * (a + b);
    }
    __syncthreads();
    // inverse iteration; groups of numerically equal eigenvalues are handled
    // by their first member, which orthogonalises within the group.
    const double ctol = 1e-10 * tnorm;
    const double tiny = 2.2e-16 * tnorm;
    for (int j = tid; j < nev; j += RB) {
        const bool leader = (j == 0) || (lam[j] - lam[j - 1] > ctol);
        if (!leader) continue;
        int gend = j + 1;
        while (gend < nev && lam[gend] - lam[gend - 1] <= ctol) ++gend;
        double* w = ws + (int64_t)j * 6 * N;
        double* y = w + 5 * N;
        for (int g = j; g < gend; ++g) {
            // start vector: deterministic, not orthogonal to anything special
            for (int i = 0; i < N; ++i) y[i] = 1.0 + 0.37 * sin(0.71 * (i + 1) * (g + 1));
            const double shift = lam[g] + (double)(g - j) * 4.0 * tiny;
            for (int it = 0; it < 3; ++it) {
                tri_solve_shifted(d, e, N, shift, tiny, y, w);
                for (int h = j; h < g; ++h) {  // orthogonalise against group
                    double s = 0.0;
                    for (int i = 0; i < N; ++i) s += y[i] * Z[(int64_t)i * nev + h];
                    for (int i = 0; i < N; ++i) y[i] -= s * Z[(int64_t)i * nev + h];
                }
                double nn = 0.0;
                for (int i = 0; i < N; ++i) nn += y[i] * y[i];
                nn = 1.0 / sqrt(fmax(nn, 1e-300));
                for (int i = 0; i < N; ++i) y[i] *= nn;
            }
            for (int i = 0; i < N; ++i) Z[(int64_t)i * nev + g] = y[i];
        }
    }
    __syncthreads();
}

// Z := Q Z with Q = H_0 ... H_{N-3} (reflectors stored below H's diagonal)
__device__ void apply_q(const double* H, const double* tau, int N, double* Z, int nev, double* dots) {
    const int tid = threadIdx.x;
    for (int k = N - 3; k >= 0; --k) {
        const double t = tau[k];
        if (t == 0.0) continue;
        const int L = N - k - 1;
        // dots[c] = v^T Z[k+1:, c]
        const int G = 16;
        const int gi = tid / G, gl = tid % G;
        for (int c = gi; c < nev; c += RB / G) {
            double s = 0.0;
            for (int i = gl; i < L; i += G) s += H[(int64_t)(k + 1 + i) * N + k] * Z[(int64_t)(k + 1 + i) * nev + c];
            for (int off = G / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
            if (gl == 0) dots[c] = t * s;
        }
        __syncthreads();
        for (int idx = tid; idx < L * nev; idx += RB) {
            int i = idx / nev, c = idx % nev;
            Z[(int64_t)(k + 1 + i) * nev + c] -= H[(int64_t)(k + 1 + i) * N + k] * dots[c];
        }
        __syncthreads();
    }
}

// Rayleigh-Ritz: GA, GM (N x N) -> C (N x nev), theta (nev).  status bit 1:
// GM not positive definite.
__global__ void __launch_bounds__(RB) k_rr(int N, int nev, double* __restrict__ GA, double* __restrict__ GM,
                                           double* __restrict__ C, double* __restrict__ theta,
                                           double* __restrict__ work, int* __restrict__ status) {
    __shared__ double red[64];
    __shared__ int flag[1];
    const int tid = threadIdx.x;
    if (tid == 0) flag[0] = 0;
    // symmetrise
    for (int idx = tid; idx < N * N; idx += RB) {
        int i = idx / N, j = idx % N;
        if (j > i) {
            double a = 0.5 * (GA[i * N + j] + GA[j * N + i]);
            GA[i * N + j] = a;
            GA[j * N + i] = a;
            double b = 0.5 * (GM[i * N + j] + GM[j * N + i]);
            GM[i * N + j] = b;
            GM[j * N + i] = b;
        }
    }
    __syncthreads();
    chol_inplace(GM, N, red, flag);
    if (flag[0]) {
        if (tid == 0) status[0] = 1;
        return;
    }
    // H = L^-1 GA L^-T : Y = L^-1 GA (in place in GA), then H = L^-1 Y^T
    trsm_lower_cols(GM, N, GA, N, N);
    double* H = work;  // N x N
    for (int idx = tid; idx < N * N; idx += RB) {
        int i = idx / N, j = idx % N;
        H[idx] = GA[j * N + i];
    }
    __syncthreads();
    trsm_lower_cols(GM, N, H, N, N);
    for (int idx = tid; idx < N * N; idx += RB) {
        int i = idx / N, j = idx % N;
        if (j > i) {
            double a = 0.5 * (H[i * N + j] + H[j * N + i]);
            H[i * N + j] = a;
            H[j * N + i] = a;
        }
    }
    __syncthreads();
    double* d = work + N * N;
    double* e = d + N;
    double* tau = e + N;
    double* v = tau + N;
    double* p = v + N;
    double* e2 = p + N;
    double* dots = e2 + N;     // nev
    double* Z = dots + N + nev;  // N x nev
    double* ws = Z + (int64_t)N * nev;  // nev * 6N
    tridiag(H, N, d, e, tau, v, p, red);
    tri_eig_lowest(d, e, N, nev, theta, Z, e2, ws, red);
    apply_q(H, tau, N, Z, nev, dots);
    // C = L^-T Z
    for (int idx = tid; idx < N * nev; idx += RB) C[idx] = Z[idx];
    __syncthreads();
    trsm_lowerT_cols(GM, N, C, nev, nev);
    if (tid == 0) status[0] = 0;
}

size_t rr_work_doubles(int N, int nev) {
    return (size_t)N * N + 7 * (size_t)N + nev + (size_t)N * nev + (size_t)nev * 6 * N + 64;
}

void rr_solve(lh_ctx* ctx, int N, int nev, double* GA, double* GM, double* C, double* theta, double* work,
              int* status) {
    k_rr<<<1, RB, 0, ctx->stream>>>(N, nev, GA, GM, C, theta, work, status);
    LH_CHECK_LAUNCH();
}

// T with T^T G T = I for a small SPD Gram G (m x m, m <= 128): G is
// equilibrated to unit diagonal (D G D), optionally shifted, and T = D U
// with U upper triangular and U^T G_s U = I, the inverse of G_s's Cholesky
// factor transposed.  U comes from symmetric elimination on [G_s | I]:
// step j takes the pivot p = S_jj of the Schur complement S (the square of
// the Cholesky factor's diagonal), subtracts r r^T / p from S (r = row j of
// S, zero at indices <= j) and U[:, j] r^T / p from U, and scales column j
// of U by 1 / sqrt(p) at the end (Gram-Schmidt of the unit vectors in the
// G_s inner product).  S and U live in registers: 512 threads form 16
// half-waves, thread (hw, ln) holding entries [hw + 16a][ln + 32b].  Columns
// are eliminated in pairs: rows j, j + 1 of S and columns j, j + 1 of U are
// published to LDS, double-buffered by the pair's parity, so one barrier per
// two columns suffices.  status 1 on a non-positive diagonal
// or a pivot <= 1e-12 (reported in ((double*)status)[1]); otherwise the
// smallest pivot is reported there (below 1e-2 the caller repeats the pass).
constexpr int CHOL_NT = 512;

template <int CB>
__global__ void __launch_bounds__(CHOL_NT) k_cholinv(int m, const double* __restrict__ G, double* __restrict__ T,
                                                     double shift, int* __restrict__ status) {
    constexpr int RA = 2 * CB;
    __shared__ double dsc[kCholMax], piv[kCholMax + 1];
    __shared__ double rbuf[2][2][32 * CB], ubuf[2][2][32 * CB];
    __shared__ int flag[1];
    const int tid = threadIdx.x, hw = tid >> 5, ln = tid & 31;
    if (tid == 0) flag[0] = 0;
    __syncthreads();
    for (int i = tid; i < m; i += CHOL_NT) {
        const double d = G[i * m + i];
        dsc[i] = d > 0.0 ? 1.0 / sqrt(d) : 0.0;
        if (!(d > 0.0)) flag[0] = 1;
    }
    __syncthreads();
    if (flag[0]) {
        if (tid == 0) status[0] = 1;
        return;
    }
    // S and U padded to 32 CB with the identity: the padding has zero
    // coupling, is never a pivot and stays unchanged, so no bounds tests
    // are needed inside the elimination
    double S[RA][CB], U[RA][CB];
#pragma unroll
    for (int a = 0; a < RA; ++a) {
        const int i = hw + 16 * a;
#pragma unroll
        for (int b = 0; b < CB; ++b) {
            const int c = ln + 32 * b;
            double v = (i == c) ? 1.0 : 0.0;
            if (i < m && c < m && i != c) v = 0.5 * (G[i * m + c] + G[c * m + i]) * dsc[i] * dsc[c];
            if (i == c && i < m) v = 1.0 + shift;
            S[a][b] = v;
            U[a][b] = (i == c) ? 1.0 : 0.0;
        }
    }
    // two columns per step (one barrier): rows j and j + 1 of S and columns j
    // and j + 1 of U are published; every thread corrects row j + 1 for the
    // elimination of column j itself (r2' = r2 - f r1, f = S[j][j+1] / p1)
    // and applies both rank-1 updates.  An odd m pairs its last column with
    // the identity padding (pivot 1, no coupling).
    for (int j = 0; j < m; j += 2) {
        const int par = (j >> 1) & 1;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int jr = j + h;
            if (hw == (jr & 15)) {
                const int aj = jr >> 4;
#pragma unroll
                for (int a = 0; a < RA; ++a)
                    if (a == aj) {
#pragma unroll
                        for (int b = 0; b < CB; ++b) {
                            const int c = ln + 32 * b;
                            rbuf[par][h][c] = (c >= jr) ? S[a][b] : 0.0;
                        }
                    }
            }
            if (ln == (jr & 31)) {
                const int bj = jr >> 5;
#pragma unroll
                for (int b = 0; b < CB; ++b)
                    if (b == bj) {
#pragma unroll
                        for (int a = 0; a < RA; ++a) ubuf[par][h][hw + 16 * a] = U[a][b];  // zero below row jr
                    }
            }
        }
        __syncthreads();
        const double p1 = rbuf[par][0][j];
        if (!(p1 > 1e-12)) {  // uniform
            if (tid == 0) {
                status[0] = 1;
                reinterpret_cast<double*>(status)[1] = p1;
            }
            return;
        }
        const double ip1 = rcp_nr(p1);  // 1e-12 < p <= ~1 (unit diagonal): the short-latency form
        const double f = rbuf[par][0][j + 1] * ip1;
        const double p2 = rbuf[par][1][j + 1] - f * rbuf[par][0][j + 1];
        if (!(p2 > 1e-12)) {  // uniform
            if (tid == 0) {
                status[0] = 1;
                reinterpret_cast<double*>(status)[1] = p2;
            }
            return;
        }
        const double ip2 = rcp_nr(p2);
        if (tid == 0) {
            piv[j] = p1;
            piv[j + 1] = p2;
        }
        double fc1[CB], fc2[CB];
#pragma unroll
        for (int b = 0; b < CB; ++b) {
            const int c = ln + 32 * b;
            const double r1 = rbuf[par][0][c];
            fc1[b] = (c > j) ? r1 * ip1 : 0.0;
            fc2[b] = (c > j + 1) ? (rbuf[par][1][c] - f * r1) * ip2 : 0.0;
        }
        // S -= r1 fc1^T + r2' fc2^T, U -= u1 fc1^T + u2' fc2^T (r1 and r2' are
        // zero at rows before j; rows j and j + 1 are finished); columns <= j
        // have zero coefficients
        const int b_lo = (j >= 31) ? (j - 31) / 32 + 1 : 0;
#pragma unroll
        for (int a = 0; a < RA; ++a) {
            const int i = hw + 16 * a;
            const double r1i = rbuf[par][0][i];
            const double r2i = (i > j) ? rbuf[par][1][i] - f * r1i : 0.0;
            const double u1i = ubuf[par][0][i];
            const double u2i = ubuf[par][1][i] - f * u1i;
#pragma unroll
            for (int b = 0; b < CB; ++b) {
                if (b < b_lo) continue;
                S[a][b] -= r1i * fc1[b] + r2i * fc2[b];
                U[a][b] -= u1i * fc1[b] + u2i * fc2[b];
            }
        }
    }
    __syncthreads();
    double isq[CB];
#pragma unroll
    for (int b = 0; b < CB; ++b) {
        const int c = ln + 32 * b;
        isq[b] = (c < m) ? 1.0 / sqrt(piv[c]) : 0.0;
    }
#pragma unroll
    for (int a = 0; a < RA; ++a) {
        const int i = hw + 16 * a;
        if (i >= m) continue;
        const double di = dsc[i];
#pragma unroll
        for (int b = 0; b < CB; ++b) {
            const int c = ln + 32 * b;
            if (c < m) T[i * m + c] = (c >= i) ? di * U[a][b] * isq[b] : 0.0;
        }
    }
    if (tid < 64) {  // the smallest pivot (a running minimum in the loop costs registers)
        double mp = 1.0;
        for (int j = tid; j < m; j += 64) mp = fmin(mp, piv[j]);
        for (int off = 32; off > 0; off >>= 1) mp = fmin(mp, __shfl_xor(mp, off, 64));
        if (tid == 0) {
            status[0] = 0;
            reinterpret_cast<double*>(status)[1] = mp;
        }
    }
}

void cholinv(lh_ctx* ctx, int m, double* G, double* T, double shift, int* status) {
    if (m > kCholMax || m < 1) LH_THROW(LH_ERR_ARG, "cholinv: block width %d outside 1..%d", m, kCholMax);
    switch ((m + 31) / 32) {
        case 1: k_cholinv<1><<<1, CHOL_NT, 0, ctx->stream>>>(m, G, T, shift, status); break;
        case 2: k_cholinv<2><<<1, CHOL_NT, 0, ctx->stream>>>(m, G, T, shift, status); break;
        case 3: k_cholinv<3><<<1, CHOL_NT, 0, ctx->stream>>>(m, G, T, shift, status); break;
        default: k_cholinv<4><<<1, CHOL_NT, 0, ctx->stream>>>(m, G, T, shift, status); break;
    }
    LH_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------
// block utilities
// ---------------------------------------------------------------------------
// R = AX - MX diag(theta), R written with ldr; X blocks with ldx
template <typename T>
__global__ void k_residual(int64_t n, int m, const T* __restrict__ AX, const T* __restrict__ MX, int ldx,
                           const double* __restrict__ theta, T* __restrict__ R, int ldr) {
    int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t i = gid / m;
    int c = (int)(gid % m);
    if (i >= n) return;
    R[i * ldr + c] = (T)((double)AX[i * ldx + c] - theta[c] * (double)MX[i * ldx + c]);
}
template <typename T>
void residual(lh_ctx* ctx, int64_t n, int m, const T* AX, const T* MX, int ldx, const double* theta, T* R, int ldr) {
    k_residual<T><<<nblocks(n * m, 256), 256, 0, ctx->stream>>>(n, m, AX, MX, ldx, theta, R, ldr);
    LH_CHECK_LAUNCH();
}
template void residual<double>(lh_ctx*, int64_t, int, const double*, const double*, int, const double*, double*, int);
template void residual<float>(lh_ctx*, int64_t, int, const float*, const float*, int, const double*, float*, int);

// per-column sums of squares of a block (n x m, ld), deterministic
template <typename T>
__global__ void k_colsq_partial(int64_t n, int m, const T* __restrict__ X, int ld, int64_t rpc,
                                double* __restrict__ part) {
    const int c = threadIdx.x % 64 + 64 * blockIdx.y;
    const int rg = threadIdx.x / 64;  // 4 row groups
    __shared__ double s[4][64];
    double acc = 0.0;
    if (c < m) {
        int64_t rb = (int64_t)blockIdx.x * rpc, re = min(rb + rpc, n);
        for (int64_t i = rb + rg; i < re; i += 4) {
            double v = (double)X[i * ld + c];
            acc += v * v;
        }
    }
    s[rg][threadIdx.x % 64] = acc;
    __syncthreads();
    if (rg == 0 && c < m) part[(int64_t)blockIdx.x * m + c] = ((s[0][threadIdx.x] + s[1][threadIdx.x]) + s[2][threadIdx.x]) + s[3][threadIdx.x];
}
// out[c] = sum over the nchunk partials of column c: one workgroup per
// column, strided per-thread sums, then a fixed-order tree (deterministic)
__global__ void __launch_bounds__(256) k_colsq_reduce(int nchunk, int m, const double* __restrict__ part,
                                                       double* __restrict__ out) {
    __shared__ double sh[256];
    const int c = blockIdx.x;
    double s = 0.0;
    for (int k = threadIdx.x; k < nchunk; k += 256) s += part[(int64_t)k * m + c];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = sh[0];
}
template <typename T>
void colsq(lh_ctx* ctx, int64_t n, int m, const T* X, int ld, double* out, DevBuf& ws) {
    int64_t nchunk = std::min<int64_t>(512, std::max<int64_t>(1, n / 256));
    int64_t rpc = (n + nchunk - 1) / nchunk;
    nchunk = (n + rpc - 1) / rpc;
    ws.ensure(nchunk * m * sizeof(double));
    dim3 grid((unsigned)nchunk, (m + 63) / 64);
    k_colsq_partial<T><<<grid, 256, 0, ctx->stream>>>(n, m, X, ld, rpc, ws.as<double>());
    LH_CHECK_LAUNCH();
    k_colsq_reduce<<<m, 256, 0, ctx->stream>>>((int)nchunk, m, ws.as<double>(), out);
    LH_CHECK_LAUNCH();
}
template void colsq<double>(lh_ctx*, int64_t, int, const double*, int, double*, DevBuf&);
template void colsq<float>(lh_ctx*, int64_t, int, const float*, int, double*, DevBuf&);

// R = AX - MX diag(theta) fused with the squared column norms of R and of
// MX (per-workgroup partials, summed in fixed order by k_colsq_reduce).
// 16-byte lanes: LPR = m / V lanes per row (V = 2 doubles or 4 floats),
// 256 / LPR rows per pass of the workgroup over its row range.
template <typename T, int LPR>
__global__ void __launch_bounds__(256)
k_residual_norms(int64_t n, int m, const T* __restrict__ AX, const T* __restrict__ MX, int ldx,
                 const double* __restrict__ theta, T* __restrict__ R, int ldr, int64_t rpc,
                 double* __restrict__ pr, double* __restrict__ pm) {
    using V = typename Vec16<T>::type;
    constexpr int VN = Vec16<T>::n;
    constexpr int RG = 256 / LPR;  // row groups
    __shared__ double sr[RG][LPR * VN], sm[RG][LPR * VN];
    const int lane = threadIdx.x % LPR, rg = threadIdx.x / LPR;
    const int c0 = lane * VN;
    double ar[VN], am[VN], th[VN];
#pragma unroll
    for (int q = 0; q < VN; ++q) {
        ar[q] = 0.0;
        am[q] = 0.0;
        th[q] = c0 + q < m ? theta[c0 + q] : 0.0;
    }
    const int64_t rb = (int64_t)blockIdx.x * rpc, re = min(rb + rpc, n);
    if (c0 < m) {
        for (int64_t i = rb + rg; i < re; i += RG) {
            const V a = *reinterpret_cast<const V*>(AX + i * ldx + c0);
            const V b = *reinterpret_cast<const V*>(MX + i * ldx + c0);
            const T* pa = reinterpret_cast<const T*>(&a);
            const T* pb = reinterpret_cast<const T*>(&b);
            V r;
            T* prr = reinterpret_cast<T*>(&r);
#pragma unroll
            for (int q = 0; q < VN; ++q) {
                const double rv = (double)pa[q] - th[q] * (double)pb[q];
                prr[q] = (T)rv;
                const double rq = (double)prr[q];
                ar[q] += rq * rq;
                am[q] += (double)pb[q] * (double)pb[q];
            }
            *reinterpret_cast<V*>(R + i * ldr + c0) = r;
        }
    }
#pragma unroll
    for (int q = 0; q < VN; ++q) {
        sr[rg][c0 + q] = ar[q];
        sm[rg][c0 + q] = am[q];
    }
    __syncthreads();
    // column c summed over the row groups in fixed order
    for (int c = threadIdx.x; c < m; c += 256) {
        double tr = 0.0, tm = 0.0;
        for (int g = 0; g < RG; ++g) {
            tr += sr[g][c];
            tm += sm[g][c];
        }
        pr[(int64_t)blockIdx.x * m + c] = tr;
        pm[(int64_t)blockIdx.x * m + c] = tm;
    }
}

template <typename T>
void residual_norms(lh_ctx* ctx, int64_t n, int m, const T* AX, const T* MX, int ldx, const double* theta, T* R,
                    int ldr, double* r2, double* mx2, DevBuf& ws) {
    constexpr int VN = Vec16<T>::n;
    auto al = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) % 16) == 0; };
    const int lanes = m / VN;
    const bool vec = m % VN == 0 && ldx % VN == 0 && ldr % VN == 0 && al(AX) && al(MX) && al(R) &&
                     (lanes == 4 || lanes == 8 || lanes == 16 || lanes == 32 || lanes == 64);
    if (!vec) {
        residual<T>(ctx, n, m, AX, MX, ldx, theta, R, ldr);
        colsq<T>(ctx, n, m, R, ldr, r2, ws);
        colsq<T>(ctx, n, m, MX, ldx, mx2, ws);
        return;
    }
    // ~2048 workgroups of whole row groups
    const int rgs = 256 / lanes;
    int64_t rpc = std::max<int64_t>(rgs, (n + 2047) / 2048);
    rpc = (rpc + rgs - 1) / rgs * rgs;
    const int64_t nchunk = (n + rpc - 1) / rpc;
    ws.ensure((size_t)nchunk * m * 2 * sizeof(double));
    double* pr = ws.as<double>();
    double* pm = pr + nchunk * m;
    switch (lanes) {
        case 4: k_residual_norms<T, 4><<<(unsigned)nchunk, 256, 0, ctx->stream>>>(n, m, AX, MX, ldx, theta, R, ldr, rpc, pr, pm); break;
        case 8: k_residual_norms<T, 8><<<(unsigned)nchunk, 256, 0, ctx->stream>>>(n, m, AX, MX, ldx, theta, R, ldr, rpc, pr, pm); break;
        case 16: k_residual_norms<T, 16><<<(unsigned)nchunk, 256, 0, ctx->stream>>>(n, m, AX, MX, ldx, theta, R, ldr, rpc, pr, pm); break;
        case 32: k_residual_norms<T, 32><<<(unsigned)nchunk, 256, 0, ctx->stream>>>(n, m, AX, MX, ldx, theta, R, ldr, rpc, pr, pm); break;
        default: k_residual_norms<T, 64><<<(unsigned)nchunk, 256, 0, ctx->stream>>>(n, m, AX, MX, ldx, theta, R, ldr, rpc, pr, pm); break;
    }
    LH_CHECK_LAUNCH();
    k_colsq_reduce<<<m, 256, 0, ctx->stream>>>((int)nchunk, m, pr, r2);
    k_colsq_reduce<<<m, 256, 0, ctx->stream>>>((int)nchunk, m, pm, mx2);
    LH_CHECK_LAUNCH();
}
template void residual_norms<double>(lh_ctx*, int64_t, int, const double*, const double*, int, const double*, double*,
                                     int, double*, double*, DevBuf&);
template void residual_norms<float>(lh_ctx*, int64_t, int, const float*, const float*, int, const double*, float*,
                                    int, double*, double*, DevBuf&);

// X[:, c] *= s[c]
template <typename T>
__global__ void k_colscale(int64_t n, int m, T* __restrict__ X, int ld, const double* __restrict__ s) {
    int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t i = gid / m;
    int c = (int)(gid % m);
    if (i >= n) return;
    X[i * ld + c] = (T)((double)X[i * ld + c] * s[c]);
}
template <typename T>
void colscale(lh_ctx* ctx, int64_t n, int m, T* X, int ld, const double* s) {
    k_colscale<T><<<nblocks(n * m, 256), 256, 0, ctx->stream>>>(n, m, X, ld, s);
    LH_CHECK_LAUNCH();
}
template void colscale<double>(lh_ctx*, int64_t, int, double*, int, const double*);
template void colscale<float>(lh_ctx*, int64_t, int, float*, int, const double*);

// X[i, c] *= dinv[i]  (Jacobi)
template <typename T>
__global__ void k_rowscale(int64_t n, int m, T* __restrict__ X, int ld, const double* __restrict__ dinv) {
    int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t i = gid / m;
    int c = (int)(gid % m);
    if (i >= n) return;
    X[i * ld + c] = (T)((double)X[i * ld + c] * dinv[i]);
}
template <typename T>
void rowscale(lh_ctx* ctx, int64_t n, int m, T* X, int ld, const double* dinv) {
    k_rowscale<T><<<nblocks(n * m, 256), 256, 0, ctx->stream>>>(n, m, X, ld, dinv);
    LH_CHECK_LAUNCH();
}
template void rowscale<double>(lh_ctx*, int64_t, int, double*, int, const double*);
template void rowscale<float>(lh_ctx*, int64_t, int, float*, int, const double*);

// dst[:, c] = src[:, perm[c]] for c < m (perm on the device)
template <typename T>
__global__ void k_gather_cols(int64_t n, int m, const T* __restrict__ src, int lds, T* __restrict__ dst, int ldd,
                              const int32_t* __restrict__ perm) {
    int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t i = gid / m;
    int c = (int)(gid % m);
    if (i >= n) return;
    dst[i * ldd + c] = src[i * lds + perm[c]];
}
template <typename T>
void gather_cols(lh_ctx* ctx, int64_t n, int m, const T* src, int lds, T* dst, int ldd, const int32_t* perm) {
    if (n * m == 0) return;
    k_gather_cols<T><<<nblocks(n * m, 256), 256, 0, ctx->stream>>>(n, m, src, lds, dst, ldd, perm);
    LH_CHECK_LAUNCH();
}
template void gather_cols<double>(lh_ctx*, int64_t, int, const double*, int, double*, int, const int32_t*);
template void gather_cols<float>(lh_ctx*, int64_t, int, const float*, int, float*, int, const int32_t*);

// copy a column block
template <typename S, typename D>
__global__ void k_copy_block(int64_t n, int m, const S* __restrict__ src, int lds, D* __restrict__ dst, int ldd) {
    int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t i = gid / m;
    int c = (int)(gid % m);
    if (i >= n) return;
    dst[i * ldd + c] = (D)src[i * lds + c];
}
template <typename S, typename D>
void copy_block(lh_ctx* ctx, int64_t n, int m, const S* src, int lds, D* dst, int ldd) {
    if (n * m == 0) return;
    k_copy_block<S, D><<<nblocks(n * m, 256), 256, 0, ctx->stream>>>(n, m, src, lds, dst, ldd);
    LH_CHECK_LAUNCH();
}
template void copy_block<double, double>(lh_ctx*, int64_t, int, const double*, int, double*, int);
template void copy_block<double, float>(lh_ctx*, int64_t, int, const double*, int, float*, int);
template void copy_block<float, double>(lh_ctx*, int64_t, int, const float*, int, double*, int);
template void copy_block<float, float>(lh_ctx*, int64_t, int, const float*, int, float*, int);

// seeded standard-normal block (splitmix64 + Box-Muller)
__device__ __forceinline__ uint64_t splitmix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
template <typename T>
__global__ void k_randn(int64_t n, int m, T* __restrict__ X, int ld, uint64_t seed) {
    int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t i = gid / m;
    int c = (int)(gid % m);
    if (i >= n) return;
    uint64_t h1 = splitmix(seed ^ (uint64_t)gid * 0x9E3779B97F4A7C15ull);
    uint64_t h2 = splitmix(h1);
    double u1 = ((h1 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    double u2 = ((h2 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    X[i * ld + c] = (T)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
}
template <typename T>
void randn(lh_ctx* ctx, int64_t n, int m, T* X, int ld, uint64_t seed) {
    k_randn<T><<<nblocks(n * m, 256), 256, 0, ctx->stream>>>(n, m, X, ld, seed);
    LH_CHECK_LAUNCH();
}
template void randn<double>(lh_ctx*, int64_t, int, double*, int, uint64_t);
template void randn<float>(lh_ctx*, int64_t, int, float*, int, uint64_t);

// dinv[i] = 1 / (diag(A)[i] - sigma * diag(B)[i])
template <typename VT>
__global__ void k_jacobi_diag(int64_t n, const int32_t* __restrict__ adiag, const VT* __restrict__ avals,
                              const int32_t* __restrict__ bdiag, const VT* __restrict__ bvals, int b_dense,
                              const VT* __restrict__ bdense, double alpha, double beta, double* __restrict__ dinv) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a = adiag[i] >= 0 ? (double)avals[adiag[i]] : 0.0;
    double b = 0.0;
    if (b_dense) b = (double)bdense[i];
    else if (bdiag && bdiag[i] >= 0) b = (double)bvals[bdiag[i]];
    double dd = alpha * b + beta * a;
    dinv[i] = (dd != 0.0) ? 1.0 / dd : 0.0;
}
void jacobi_diag(lh_ctx* ctx, const lh_csr* A, const lh_csr* B, double alpha, double beta, double* dinv) {
    const int64_t n = A->n();
    if (A->dtype == LH_F64)
        k_jacobi_diag<double><<<nblocks(n, 256), 256, 0, ctx->stream>>>(
            n, A->pat->diagpos.as<int32_t>(), A->vals.as<double>(), B ? B->pat->diagpos.as<int32_t>() : nullptr,
            B ? B->vals.as<double>() : nullptr, B && B->is_diag, B && B->is_diag ? B->dense_diag.as<double>() : nullptr,
            alpha, beta, dinv);
    else
        k_jacobi_diag<float><<<nblocks(n, 256), 256, 0, ctx->stream>>>(
            n, A->pat->diagpos.as<int32_t>(), A->vals.as<float>(), B ? B->pat->diagpos.as<int32_t>() : nullptr,
            B ? B->vals.as<float>() : nullptr, B && B->is_diag, B && B->is_diag ? B->dense_diag.as<float>() : nullptr,
            alpha, beta, dinv);
    LH_CHECK_LAUNCH();
}

}  // namespace lh

using namespace lh;

extern "C" int lh_small_cholinv(lh_ctx* ctx, int m, const double* G, double shift, double* T, double* min_pivot) {
    if (!ctx || !G || !T || m < 1 || m > kCholMax || !(shift >= 0.0)) return LH_ERR_ARG;
    LH_API_BEGIN(ctx)
    HIPX(hipSetDevice(ctx->device));
    DevBuf dG((size_t)m * m * 8), dT((size_t)m * m * 8), st(64);
    h2d(ctx, dG.p, G, (size_t)m * m * 8);
    HIPX(hipMemsetAsync(st.p, 0, 64, ctx->stream));
    cholinv(ctx, m, dG.as<double>(), dT.as<double>(), shift, st.as<int>());
    double sv[2];
    d2h(ctx, sv, st.p, sizeof(sv));
    if (min_pivot) *min_pivot = sv[1];
    if (reinterpret_cast<int*>(sv)[0])
        LH_THROW(LH_ERR_NUMERIC, "lh_small_cholinv: not numerically positive definite (pivot %g)", sv[1]);
    d2h(ctx, T, dT.p, (size_t)m * m * 8);
    LH_API_END(ctx)
}
