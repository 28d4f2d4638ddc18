// Rayleigh-Ritz eigensolver for an orthonormal trial basis: the smallest nev
// eigenpairs of a small dense symmetric H (N <= 192), FP64, one workgroup of
// 1024 threads, working set in LDS.
//
//  1. the lower triangle of H is held packed in LDS (N(N+1)/2 doubles,
//     148 KB at N = 192) and reduced to tridiagonal form by Householder
//     reflections; every row of the trailing matrix is owned by four lanes,
//     so the symmetric matvec and the rank-2 update walk contiguous packed
//     rows without index arithmetic in the inner loops;
//  2. the reflectors are parked in global memory, freeing LDS for Z;
//  3. the wanted eigenvalues of the tridiagonal T are found by multisection:
//     16 lanes per eigenvalue evaluate Sturm counts at 16 interior points,
//     shrinking the bracket 17-fold per round;
//  4. their eigenvectors by inverse iteration (LU with partial pivoting of
//     T - lambda I, per-thread scratch in global memory); numerically equal
//     eigenvalues are handled by one thread with Gram-Schmidt inside the group;
//  5. Z := Q Z applies the reflectors in reverse order with Z in LDS, the next
//     reflector prefetched into registers while the current one is applied;
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

__device__ __forceinline__ int rowbase(int i) { return i * (i + 1) / 2; }

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
#pragma unroll
    for (int i = 0; i < NB / 64; ++i) s += red[i];
    return s;
}

__device__ __forceinline__ void stamp(unsigned long long* prof, int slot) {
    if (prof && threadIdx.x == 0) prof[slot] = wall_clock64();
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
    double a = y[0];
    for (int i = 0; i < N - 1; ++i) {
        double b = y[(i + 1) * ys];
        if (pv[i] != 0.0) { double t = a; a = b; b = t; }
        y[i * ys] = a;
        a = b - lm[i] * a;
    }
    double x2 = a / u0[N - 1];
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
         double* __restrict__ work, int* __restrict__ status, unsigned long long* __restrict__ prof) {
    __shared__ double H[PK];  // packed lower triangle of H; later Z (N x nev) and the small Gram
    __shared__ double sv[NMAX], sp[NMAX], sd[NMAX], se[NMAX], se2[NMAX], stau[NMAX];
    __shared__ double red[32];
    const int tid = threadIdx.x;
    const int rl = tid >> 2, gl = tid & 3;  // row / lane-in-row for the 4-lanes-per-row loops
    stamp(prof, 0);

    // 1. load (symmetrised) packed lower triangle, 4 lanes per row
    for (int i = rl; i < N; i += NB / 4) {
        const int b = rowbase(i);
        for (int j = gl; j <= i; j += 4) H[b + j] = 0.5 * (G[(int64_t)i * ldg + j] + G[(int64_t)j * ldg + i]);
    }
    __syncthreads();
    stamp(prof, 1);

    // 2. Householder tridiagonalisation
    for (int k = 0; k < N - 2; ++k) {
        const int L = N - k - 1;
        double part = 0.0;
        for (int i = tid; i < L; i += NB) {
            const double x = H[rowbase(k + 1 + i) + k];
            sv[i] = x;
            part += x * x;
        }
        const double xn2 = bsum(part, red);
        const double x0 = sv[0];
        if (xn2 == 0.0) {
            if (tid == 0) { sd[k] = H[rowbase(k) + k]; se[k] = 0.0; stau[k] = 0.0; }
            __syncthreads();
            continue;
        }
        const double alpha = (x0 >= 0.0) ? -sqrt(xn2) : sqrt(xn2);
        const double v0 = x0 - alpha;
        const double t = 2.0 / (xn2 - x0 * x0 + v0 * v0);
        __syncthreads();
        if (tid == 0) {
            sv[0] = v0;
            H[rowbase(k + 1) + k] = v0;
            sd[k] = H[rowbase(k) + k];
            se[k] = alpha;
            stau[k] = t;
        }
        __syncthreads();
        // p = t * H22 v ; row r of H22 = packed row i = k+1+r (columns <= i
        // contiguous) plus column i below it (entries j > i, stride j)
        for (int r = rl; r < L; r += NB / 4) {
            const int i = k + 1 + r;
            const double* hr = H + rowbase(i) + (k + 1);
            double s = 0.0;
            for (int c = gl; c <= r; c += 4) s += hr[c] * sv[c];
            for (int c = r + 1 + gl; c < L; c += 4) s += H[rowbase(k + 1 + c) + i] * sv[c];
            s += __shfl_xor(s, 1, 64);
            s += __shfl_xor(s, 2, 64);
            if (gl == 0) sp[r] = t * s;
        }
        __syncthreads();
        double pp = 0.0;
        for (int i = tid; i < L; i += NB) pp += sp[i] * sv[i];
        const double K = 0.5 * t * bsum(pp, red);
        for (int i = tid; i < L; i += NB) sp[i] -= K * sv[i];
        __syncthreads();
        for (int r = rl; r < L; r += NB / 4) {
            double* hr = H + rowbase(k + 1 + r) + (k + 1);
            const double vr = sv[r], wr = sp[r];
            for (int c = gl; c <= r; c += 4) hr[c] -= vr * sp[c] + wr * sv[c];
        }
        __syncthreads();
    }
    if (tid == 0) {
        if (N >= 2) {
            sd[N - 2] = H[rowbase(N - 2) + N - 2];
            se[N - 2] = H[rowbase(N - 1) + N - 2];
            stau[N - 2] = 0.0;
        }
        sd[N - 1] = H[rowbase(N - 1) + N - 1];
        stau[N - 1] = 0.0;
    }
    stamp(prof, 2);
    // 3. park reflectors in global memory: R[k][i] = v_k[i]
    double* R = work;
    for (int k = rl; k < N - 2; k += NB / 4)
        for (int i = gl; i < N - k - 1; i += 4) R[(int64_t)k * N + i] = H[rowbase(k + 1 + i) + k];
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
    stamp(prof, 3);
    // 5. inverse iteration; Z (N x nev) lives in LDS (H is free now).  The
    // eigenvalue threads are spread over the waves (4 per wave) so that the
    // latency of the scalar recurrences is hidden across waves.
    double* Z = H;
    const double tiny = 2.2e-16 * tnorm;
    {
        const int w = tid >> 6, l = tid & 63;
        const int j = (l < 4) ? (w * 4 + l) : -1;  // 64 eigenvalue slots per pass
        for (int jb = 0; jb < nev; jb += 64) {
            const int jj = (j >= 0) ? jb + j : nev;
            if (jj < nev) {
                const double scale_j = fmax(fabs(lam[jj]), 1e-6 * tnorm);
                const bool leader = (jj == 0) || (lam[jj] - lam[jj - 1] > 1e-7 * scale_j);
                if (leader) {
                    int gend = jj + 1;
                    while (gend < nev && lam[gend] - lam[gend - 1] <= 1e-7 * fmax(fabs(lam[gend]), 1e-6 * tnorm))
                        ++gend;
                    double* ws = work + (int64_t)N * N + (int64_t)jj * 5 * N;
                    for (int g = jj; g < gend; ++g) {
                        double* y = Z + g;  // column g, stride nev
                        for (int i = 0; i < N; ++i) y[i * nev] = 1.0 + 0.1 * (double)((7 * i + 3 * g) % 11);
                        tri_factor(sd, se, N, lam[g] + (double)(g - jj) * 8.0 * tiny, tiny, ws);
                        for (int itr = 0; itr < 2; ++itr) {
                            tri_lu_solve(ws, N, y, nev);
                            for (int h = jj; h < g; ++h) {
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
            }
        }
    }
    __syncthreads();
    for (int j = tid; j < nev; j += NB) theta[j] = lam[j];
    stamp(prof, 4);
    // 6. Z := Q Z, reflectors k = N-3 .. 0; sv holds the current reflector,
    // the next one is loaded into a register before the current is applied
    {
        int k = N - 3;
        double nxt = (k >= 0 && tid < N - k - 1) ? R[(int64_t)k * N + tid] : 0.0;
        for (; k >= 0; --k) {
            const int L = N - k - 1;
            if (tid < L) sv[tid] = nxt;
            __syncthreads();
            nxt = (k - 1 >= 0 && tid < N - k) ? R[(int64_t)(k - 1) * N + tid] : 0.0;
            const double t = stau[k];
            if (t != 0.0) {
                for (int c0 = 0; c0 < nev; c0 += NB / 16) {
                    const int c = c0 + tid / 16, g16 = tid % 16;
                    double s = 0.0;
                    if (c < nev)
                        for (int i = g16; i < L; i += 16) s += sv[i] * Z[(k + 1 + i) * nev + c];
                    for (int off = 8; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
                    if (c < nev && g16 == 0) se2[c] = t * s;  // se2 reused as dots
                }
                __syncthreads();
                for (int i = rl; i < L; i += NB / 4) {
                    double* zr = Z + (k + 1 + i) * nev;
                    const double vi = sv[i];
                    for (int c = gl; c < nev; c += 4) zr[c] -= vi * se2[c];
                }
            }
            __syncthreads();
        }
    }
    stamp(prof, 5);
    // 7. Cholesky-QR of Z in the small space (when the Gram fits in LDS)
    if (N * nev + nev * nev <= PK) {
        double* Gs = H + N * nev;
        for (int idx = tid; idx < nev * nev; idx += NB) {
            const int a = idx / nev, b = idx - a * nev;
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
            const int Lr = nev - jj - 1;
            // column scale and trailing update fused: lanes own rows
            for (int i = jj + 1 + (tid >> 4); i < nev; i += NB / 16) {
                const double lij = Gs[i * nev + jj] / ljj;
                // row i's trailing entries kk in (jj, i]; column jj of row kk
                // is scaled by the lane owning row kk in this same pass, so
                // read the unscaled value and divide here
                for (int kk = jj + 1 + (tid & 15); kk <= i; kk += 16)
                    Gs[i * nev + kk] -= lij * (Gs[kk * nev + jj] / ljj);
            }
            __syncthreads();
            for (int i = jj + 1 + tid; i < nev; i += NB) Gs[i * nev + jj] /= ljj;
            __syncthreads();
            (void)Lr;
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
    stamp(prof, 6);
    for (int idx = tid; idx < N * nev; idx += NB) C[idx] = Z[idx];
    if (tid == 0) status[0] = 0;
    stamp(prof, 7);
}

}  // namespace

size_t rr_std_work_doubles(int N, int nev) { return (size_t)N * N + (size_t)nev * 5 * N + 64; }

void rr_std(lh_ctx* ctx, int N, int nev, const double* G, int ldg, double* C, double* theta, double* work,
            int* status, unsigned long long* prof) {
    if (N > kRRStdMaxN || nev > N || N < 1) LH_THROW(LH_ERR_ARG, "rr_std: N=%d nev=%d out of range", N, nev);
    k_rr_std<<<1, NB, 0, ctx->stream>>>(N, nev, G, ldg, C, theta, work, status, prof);
    LH_CHECK_LAUNCH();
}

}  // namespace lh

using namespace lh;

extern "C" int lh_small_symeig(lh_ctx* ctx, int N, int nev, const double* H, double* evals, double* Z,
                               double* phase_us) {
    if (!ctx || !H || N < 1 || nev < 1 || nev > N || N > kRRStdMaxN) return LH_ERR_ARG;
    LH_API_BEGIN(ctx)
    HIPX(hipSetDevice(ctx->device));
    DevBuf dH((size_t)N * N * 8), dC((size_t)N * nev * 8), dth((size_t)nev * 8),
        work(rr_std_work_doubles(N, nev) * 8), st(64), prof(8 * 8);
    h2d(ctx, dH.p, H, (size_t)N * N * 8);
    HIPX(hipMemsetAsync(prof.p, 0, 64, ctx->stream));
    rr_std(ctx, N, nev, dH.as<double>(), N, dC.as<double>(), dth.as<double>(), work.as<double>(), st.as<int>(),
           phase_us ? prof.as<unsigned long long>() : nullptr);
    if (evals) d2h(ctx, evals, dth.p, (size_t)nev * 8);
    if (Z) d2h(ctx, Z, dC.p, (size_t)N * nev * 8);  // row-major N x nev
    if (phase_us) {
        unsigned long long p[8];
        d2h(ctx, p, prof.p, sizeof(p));
        int khz = 100000;
        (void)hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device);
        for (int i = 0; i < 7; ++i) phase_us[i] = (double)(p[i + 1] - p[i]) * 1000.0 / (double)khz;
    }
    HIPX(hipStreamSynchronize(ctx->stream));
    LH_API_END(ctx)
}
