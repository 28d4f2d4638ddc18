// Rayleigh-Ritz eigensolver for an orthonormal trial basis: the smallest nev
// eigenpairs of a small dense symmetric H (N <= 192), FP64, one workgroup of
// 1024 threads, working set in LDS.
//
//  1. the lower triangle of H is held packed in LDS (N(N+1)/2 doubles,
//     148 KB at N = 192) and reduced to tridiagonal form by Householder
//     reflections (two-sided, one reflector per column);
//  2. the reflectors are parked in global memory, freeing LDS for Z;
//  3. the wanted eigenvalues of the tridiagonal T are found by multisection:
//     16 lanes per eigenvalue evaluate Sturm counts at 16 interior points,
//     shrinking the bracket 17-fold per round;
//  4. their eigenvectors by inverse iteration (LU with partial pivoting of
//     T - lambda I, per-thread scratch in global memory); numerically equal
//     eigenvalues are handled by one thread with Gram-Schmidt inside the group;
//  5. Z := Q Z applies the reflectors in reverse order, Z in LDS;
//  6. a Cholesky-QR of Z in the small space restores orthonormality to
//     rounding level (inverse iteration leaves errors ~ eps ||T|| / gap).
//
// Replaces the dense Rayleigh-Ritz step ARPACK performs inside eigsh
// (lapy/solver.py:713-715) for the block solver.
#include "lh_dense.hpp"

#include <cmath>

namespace lh {

namespace {

constexpr int NB = 1024;
constexpr int NMAX = kRRStdMaxN;
constexpr int PK = NMAX * (NMAX + 1) / 2;

__device__ __forceinline__ int pidx(int i, int j) { return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }

__device__ __forceinline__ double wsum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ double bsum(double v, double* red) {
    v = wsum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < NB / 64; ++i) s += red[i];
    return s;
}

__device__ int sturm(const double* d, const double* e2, int N, double x, double pivmin) {
    int cnt = 0;
    double q = d[0] - x;
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += q < 0.0;
    for (int i = 1; i < N; ++i) {
        q = (d[i] - x) - e2[i - 1] / q;
        if (fabs(q) < pivmin) q = -pivmin;
        cnt += q < 0.0;
    }
    return cnt;
}

// LU with partial pivoting of T - lam I into ws (u0, u1, u2, l, piv: 5N)
__device__ void tri_factor(const double* d, const double* e, int N, double lam, double tiny, double* ws) {
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
            const double piv = (c0 != 0.0) ? c0 : tiny;
            const double l = sub / piv;
            u0[i] = piv; u1[i] = c1; u2[i] = c2; lm[i] = l; pv[i] = 0.0;
            c0 = a1 - l * c1;
            c1 = b1 - l * c2;
        } else {
            const double l = c0 / sub;
            u0[i] = sub; u1[i] = a1; u2[i] = b1; lm[i] = l; pv[i] = 1.0;
            c0 = c1 - l * a1;
            c1 = c2 - l * b1;
        }
        c2 = 0.0;
    }
    u0[N - 1] = (fabs(c0) > tiny) ? c0 : tiny;
}

// solve with the factorisation; y (stride ys) overwritten
__device__ void tri_lu_solve(const double* ws, int N, double* y, int ys) {
    const double* u0 = ws;
    const double* u1 = ws + N;
    const double* u2 = ws + 2 * N;
    const double* lm = ws + 3 * N;
    const double* pv = ws + 4 * N;
    for (int i = 0; i < N - 1; ++i) {
        double a = y[i * ys], b = y[(i + 1) * ys];
        if (pv[i] != 0.0) { double t = a; a = b; b = t; }
        y[i * ys] = a;
        y[(i + 1) * ys] = b - lm[i] * a;
    }
    double x2 = y[(N - 1) * ys] / u0[N - 1];
    y[(N - 1) * ys] = x2;
    if (N >= 2) {
        double x1 = (y[(N - 2) * ys] - u1[N - 2] * x2) / u0[N - 2];
        y[(N - 2) * ys] = x1;
        for (int i = N - 3; i >= 0; --i) {
            double x0 = (y[i * ys] - u1[i] * x1 - u2[i] * x2) / u0[i];
            y[i * ys] = x0;
            x2 = x1;
            x1 = x0;
        }
    }
}

__global__ void __launch_bounds__(NB)
k_rr_std(int N, int nev, const double* __restrict__ G, int ldg, double* __restrict__ C, double* __restrict__ theta,
         double* __restrict__ work, int* __restrict__ status) {
    __shared__ double H[PK];  // packed lower triangle of H; later Z (N x nev) and the small Gram
    __shared__ double sv[NMAX], sp[NMAX], sd[NMAX], se[NMAX], se2[NMAX], stau[NMAX];
    __shared__ double red[32];
    const int tid = threadIdx.x;

    // 1. load (symmetrised) packed lower triangle
    for (int idx = tid; idx < N * N; idx += NB) {
        const int i = idx / N, j = idx % N;
        if (j <= i) H[i * (i + 1) / 2 + j] = 0.5 * (G[(int64_t)i * ldg + j] + G[(int64_t)j * ldg + i]);
    }
    __syncthreads();

    // 2. Householder tridiagonalisation
    for (int k = 0; k < N - 2; ++k) {
        const int L = N - k - 1;
        double part = 0.0;
        for (int i = tid; i < L; i += NB) {
            const double x = H[pidx(k + 1 + i, k)];
            sv[i] = x;
            part += x * x;
        }
        const double xn2 = bsum(part, red);
        const double x0 = sv[0];
        if (xn2 == 0.0) {
            if (tid == 0) { sd[k] = H[pidx(k, k)]; se[k] = 0.0; stau[k] = 0.0; }
            __syncthreads();
            continue;
        }
        const double alpha = (x0 >= 0.0) ? -sqrt(xn2) : sqrt(xn2);
        const double v0 = x0 - alpha;
        const double t = 2.0 / (xn2 - x0 * x0 + v0 * v0);
        __syncthreads();
        if (tid == 0) {
            sv[0] = v0;
            H[pidx(k + 1, k)] = v0;
            sd[k] = H[pidx(k, k)];
            se[k] = alpha;
            stau[k] = t;
        }
        __syncthreads();
        // p = t * H22 v, four lanes per row
        {
            const int r = tid >> 2, gl = tid & 3;
            double s = 0.0;
            if (r < L) {
                const int i = k + 1 + r;
                for (int c = gl; c < L; c += 4) s += H[pidx(i, k + 1 + c)] * sv[c];
            }
            s += __shfl_xor(s, 1, 64);
            s += __shfl_xor(s, 2, 64);
            if (r < L && gl == 0) sp[r] = t * s;
        }
        __syncthreads();
        double pp = 0.0;
        for (int i = tid; i < L; i += NB) pp += sp[i] * sv[i];
        const double K = 0.5 * t * bsum(pp, red);
        for (int i = tid; i < L; i += NB) sp[i] -= K * sv[i];
        __syncthreads();
        for (int idx = tid; idx < L * L; idx += NB) {
            const int r = idx / L, c = idx - r * L;
            if (c <= r) {
                const int i = k + 1 + r, j = k + 1 + c;
                H[i * (i + 1) / 2 + j] -= sv[r] * sp[c] + sp[r] * sv[c];
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        if (N >= 2) {
            sd[N - 2] = H[pidx(N - 2, N - 2)];
            se[N - 2] = H[pidx(N - 1, N - 2)];
            stau[N - 2] = 0.0;
        }
        sd[N - 1] = H[pidx(N - 1, N - 1)];
        if (N >= 1) stau[N - 1] = 0.0;
    }
    // 3. park reflectors in global memory: R[k][i] = v_k[i]
    double* R = work;
    for (int idx = tid; idx < N * N; idx += NB) {
        const int k = idx / N, i = idx % N;
        if (k < N - 2 && i < N - k - 1) R[idx] = H[pidx(k + 1 + i, k)];
    }
    __syncthreads();
    for (int i = tid; i < N - 1; i += NB) se2[i] = se[i] * se[i];
    if (tid == 0) {
        double lo = sd[0], hi = sd[0], nrm = 0.0;
        for (int i = 0; i < N; ++i) {
            const double rr = (i > 0 ? fabs(se[i - 1]) : 0.0) + (i < N - 1 ? fabs(se[i]) : 0.0);
            lo = fmin(lo, sd[i] - rr);
            hi = fmax(hi, sd[i] + rr);
            nrm = fmax(nrm, fabs(sd[i]) + rr);
        }
        red[0] = lo;
        red[1] = hi;
        red[2] = fmax(nrm, 1e-300);
    }
    __syncthreads();
    const double glo = red[0], ghi = red[1], tnorm = red[2];
    const double pivmin = 1e-290 + 1e-300 * tnorm;
    // 4. multisection: 16 lanes per eigenvalue, 64 eigenvalues per pass
    double* lam = sp;  // reuse
    for (int base = 0; base < nev; base += NB / 16) {
        const int j = base + tid / 16, q = tid % 16;
        double a = glo - 1e-14 * tnorm, b = ghi + 1e-14 * tnorm;
        for (int it = 0; it < 40; ++it) {
            const double h = (b - a) / 17.0;
            const double x = a + (q + 1) * h;
            const bool pred = (j < nev) && (sturm(sd, se2, N, x, pivmin) > j);
            const unsigned long long bal = __ballot(pred);
            const unsigned grp = (unsigned)((bal >> (16 * ((tid & 63) >> 4))) & 0xFFFFull);
            const int first = grp ? (__ffs(grp) - 1) : 16;
            const double nb = (first < 16) ? a + (first + 1) * h : b;
            const double na = (first > 0) ? a + first * h : a;
            a = na;
            b = nb;
            if (b - a <= 2.2e-16 * (fabs(a) + fabs(b)) + 1e-17 * tnorm) break;
        }
        if (j < nev && q == 0) lam[j] = 0.5 * (a + b);
    }
    __syncthreads();
    // 5. inverse iteration; Z (N x nev) lives in LDS (H is free now)
    double* Z = H;
    const double tiny = 2.2e-16 * tnorm;
    for (int j = tid; j < nev; j += NB) {
        const double scale_j = fmax(fabs(lam[j]), 1e-6 * tnorm);
        const bool leader = (j == 0) || (lam[j] - lam[j - 1] > 1e-7 * scale_j);
        if (!leader) continue;
        int gend = j + 1;
        while (gend < nev && lam[gend] - lam[gend - 1] <= 1e-7 * fmax(fabs(lam[gend]), 1e-6 * tnorm)) ++gend;
        double* ws = work + (int64_t)N * N + (int64_t)j * 5 * N;
        for (int g = j; g < gend; ++g) {
            double* y = Z + g;  // column g, stride nev
            for (int i = 0; i < N; ++i) y[i * nev] = 1.0 + 0.1 * (double)((7 * i + 3 * g) % 11);
            tri_factor(sd, se, N, lam[g] + (double)(g - j) * 8.0 * tiny, tiny, ws);
            for (int itr = 0; itr < 2; ++itr) {
                tri_lu_solve(ws, N, y, nev);
                for (int h = j; h < g; ++h) {
                    double s = 0.0;
                    for (int i = 0; i < N; ++i) s += y[i * nev] * Z[i * nev + h];
                    for (int i = 0; i < N; ++i) y[i * nev] -= s * Z[i * nev + h];
                }
                double nn = 0.0;
                for (int i = 0; i < N; ++i) nn += y[i * nev] * y[i * nev];
                nn = 1.0 / sqrt(fmax(nn, 1e-300));
                for (int i = 0; i < N; ++i) y[i * nev] *= nn;
            }
        }
    }
    __syncthreads();
    if (tid < nev) theta[tid] = lam[tid];
    for (int j = tid; j < nev; j += NB) theta[j] = lam[j];
    // 6. Z := Q Z, reflectors k = N-3 .. 0
    for (int k = N - 3; k >= 0; --k) {
        const double t = stau[k];
        if (t == 0.0) continue;
        const int L = N - k - 1;
        for (int i = tid; i < L; i += NB) sv[i] = R[(int64_t)k * N + i];
        __syncthreads();
        for (int c0 = 0; c0 < nev; c0 += NB / 16) {
            const int c = c0 + tid / 16, gl = tid % 16;
            double s = 0.0;
            if (c < nev)
                for (int i = gl; i < L; i += 16) s += sv[i] * Z[(k + 1 + i) * nev + c];
            for (int off = 8; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
            if (c < nev && gl == 0) se2[c] = t * s;  // se2 reused as dots
        }
        __syncthreads();
        for (int idx = tid; idx < L * nev; idx += NB) {
            const int i = idx / nev, c = idx - i * nev;
            Z[(k + 1 + i) * nev + c] -= sv[i] * se2[c];
        }
        __syncthreads();
    }
    // 7. Cholesky-QR of Z in the small space (when the Gram fits in LDS)
    if (N * nev + nev * nev <= PK) {
        double* Gs = H + N * nev;
        for (int idx = tid; idx < nev * nev; idx += NB) {
            const int a = idx / nev, b = idx % nev;
            double s = 0.0;
            if (b <= a)
                for (int i = 0; i < N; ++i) s += Z[i * nev + a] * Z[i * nev + b];
            Gs[idx] = s;
        }
        __syncthreads();
        bool ok = true;
        for (int jj = 0; jj < nev; ++jj) {
            if (tid == 0) {
                const double dd = Gs[jj * nev + jj];
                Gs[jj * nev + jj] = dd > 0.5 ? sqrt(dd) : -1.0;
            }
            __syncthreads();
            const double ljj = Gs[jj * nev + jj];
            if (ljj < 0.0) { ok = false; break; }
            for (int i = jj + 1 + tid; i < nev; i += NB) Gs[i * nev + jj] /= ljj;
            __syncthreads();
            const int Lr = nev - jj - 1;
            for (int idx = tid; idx < Lr * Lr; idx += NB) {
                const int i = jj + 1 + idx / Lr, kk = jj + 1 + idx % Lr;
                if (kk <= i) Gs[i * nev + kk] -= Gs[i * nev + jj] * Gs[kk * nev + jj];
            }
            __syncthreads();
        }
        if (ok) {
            // each row z of Z: solve z_new L^T = z
            for (int i = tid; i < N; i += NB) {
                double* z = Z + i * nev;
                for (int c = 0; c < nev; ++c) {
                    double s = z[c];
                    for (int h = 0; h < c; ++h) s -= Gs[c * nev + h] * z[h];
                    z[c] = s / Gs[c * nev + c];
                }
            }
            __syncthreads();
        }
    }
    for (int idx = tid; idx < N * nev; idx += NB) C[idx] = Z[idx];
    if (tid == 0) status[0] = 0;
}

}  // namespace

size_t rr_std_work_doubles(int N, int nev) { return (size_t)N * N + (size_t)nev * 5 * N + 64; }

void rr_std(lh_ctx* ctx, int N, int nev, const double* G, int ldg, double* C, double* theta, double* work,
            int* status) {
    if (N > kRRStdMaxN || nev > N) LH_THROW(LH_ERR_ARG, "rr_std: N=%d nev=%d out of range", N, nev);
    k_rr_std<<<1, NB, 0, ctx->stream>>>(N, nev, G, ldg, C, theta, work, status);
    LH_CHECK_LAUNCH();
}

}  // namespace lh
