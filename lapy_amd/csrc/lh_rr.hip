// Rayleigh-Ritz eigensolver for an orthonormal trial basis: the smallest nev
// eigenpairs of a small dense symmetric H (N <= 192), FP64.
//
// Five kernels on the library stream, each bound by the latency of its
// dependent steps on one workgroup (or one wave per eigenpair):
//  1. k_rr_tri (one workgroup, 512 threads): H is held in registers, each
//     row spread over the 16 lanes of one DPP row, and reduced to
//     tridiagonal form by Householder reflections, two barriers per step.
//     The reflectors, the tridiagonal T and its Gershgorin bounds are parked
//     in global memory.
//  2. k_rr_bisect (one wave per wanted eigenvalue): multisection on T scaled
//     by a power of two to ||T|| <= 1, the 64 lanes evaluating Sturm counts
//     at 64 interior points (65-fold bracket reduction per round).
//  3. k_rr_invit (one wave per eigenvalue; only the first of a group of
//     numerically equal eigenvalues works): inverse iteration with the LU
//     factors of T - lambda I (same scaling) in LDS, Gram-Schmidt inside the
//     group.
//  4. k_rr_back (one wave per eigenvector): Z := Q Y, the reflectors applied
//     in reverse order, loaded four steps ahead into registers.
//  5. k_rr_polish (one workgroup): Newton-Schulz orthonormalisation of Z in
//     the small space, Z := Z (3 I - Z^T Z) / 2 (inverse iteration leaves
//     orthogonality errors ~ eps ||T|| / gap; each step squares them; up to
//     three steps, none for a basis with ||Z^T Z - I|| >= 1/2), its products
//     in 4 x 4 register tiles from 32 columns up.
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

__device__ __forceinline__ double wsum(double v) { return wave_sum64(v); }
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
    if (prof && threadIdx.x == 0 && blockIdx.x == 0) prof[slot] = wall_clock64();
}

// work layout (doubles): reflectors R[k][i] (N x N), d (N), e (N), tau (N),
// bounds (4: lo, hi, ||T||, -), lam (nev), Y (nev x N, column g contiguous),
// Z (N x nev row-major)
struct RRWork {
    double *R, *d, *e, *tau, *bnd, *lam, *Y, *Z;
    __host__ __device__ RRWork(double* w, int N, int nev) {
        R = w;
        d = R + (size_t)N * N;
        e = d + N;
        tau = e + N;
        bnd = tau + N;
        lam = bnd + 4;
        Y = lam + nev;
        Z = Y + (size_t)N * nev;
    }
};

// number of eigenvalues of T below x (negative pivots of T - x I).  The
// caller scales T so that ||T|| <= 1: the pivots then lie between pivmin and
// ~1 / pivmin, where the short-latency reciprocal is accurate
__device__ int sturm(const double* d, const double* e2, int N, double x, double pivmin) {
    int cnt = 0;
    double q = d[0] - x;
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += q < 0.0;
    for (int i = 1; i < N; ++i) {
        q = (d[i] - x) - e2[i - 1] * rcp_nr(q);
        if (fabs(q) < pivmin) q = -pivmin;
        cnt += q < 0.0;
    }
    return cnt;
}

// LU with partial pivoting of T - lam I into ws (1/u0, u1, u2, l, piv: 5N).
// The caller scales T to ||T|| <= 1, so the pivots (a zero or sub-normal one
// replaced by tiny) have ordinary magnitudes and their reciprocals come from
// the short-latency form, stored for the solves
__device__ void tri_factor(const double* d, const double* e, int N, double lam, double tiny, double* ws) {
    double* iu0 = ws;
    double* u1 = ws + N;
    double* u2 = ws + 2 * N;
    double* lm = ws + 3 * N;
    double* pv = ws + 4 * N;
    double c0 = d[0] - lam, c1 = (N > 1) ? e[0] : 0.0, c2 = 0.0;
    for (int i = 0; i < N - 1; ++i) {
        const double sub = e[i];
        const double a1 = d[i + 1] - lam;
        const double b1 = (i + 1 < N - 1) ? e[i + 1] : 0.0;
        if (fabs(c0) >= fabs(sub) || !(fabs(sub) > 1e-300)) {
            const double piv = (fabs(c0) > 1e-300) ? c0 : tiny;
            const double ip = rcp_nr(piv);
            const double l = sub * ip;
            iu0[i] = ip; u1[i] = c1; u2[i] = c2; lm[i] = l; pv[i] = 0.0;
            c0 = a1 - l * c1;
            c1 = b1 - l * c2;
        } else {
            const double ip = rcp_nr(sub);
            const double l = c0 * ip;
            iu0[i] = ip; u1[i] = a1; u2[i] = b1; lm[i] = l; pv[i] = 1.0;
            c0 = c1 - l * a1;
            c1 = c2 - l * b1;
        }
        c2 = 0.0;
    }
    iu0[N - 1] = rcp_nr((fabs(c0) > tiny) ? c0 : tiny);
}

// solve with the factorisation; y overwritten
__device__ void tri_lu_solve(const double* ws, int N, double* y) {
    const double* iu0 = ws;
    const double* u1 = ws + N;
    const double* u2 = ws + 2 * N;
    const double* lm = ws + 3 * N;
    const double* pv = ws + 4 * N;
    double a = y[0];
    for (int i = 0; i < N - 1; ++i) {
        double b = y[i + 1];
        if (pv[i] != 0.0) { double t = a; a = b; b = t; }
        y[i] = a;
        a = b - lm[i] * a;
    }
    double x2 = a * iu0[N - 1];
    y[N - 1] = x2;
    if (N >= 2) {
        double x1 = (y[N - 2] - u1[N - 2] * x2) * iu0[N - 2];
        y[N - 2] = x1;
        for (int i = N - 3; i >= 0; --i) {
            double x0 = (y[i] - u1[i] * x1 - u2[i] * x2) * iu0[i];
            y[i] = x0;
            x2 = x1;
            x1 = x0;
        }
    }
}

// ---- 1. tridiagonalisation ------------------------------------------------
// H lives in registers: 512 threads form 32 row groups of 16 lanes (one DPP
// row each), and the thread (group g, lane l) holds H[g + 32a][l + 16b] for
// a < CB, b < 2 CB (CB = ceil(N / 32)), a block-cyclic layout in which every
// row is spread over the 16 lanes of one group, so that row sums are DPP
// sums within a row.  Step k (Householder reflector of row k = column k
// below the diagonal) costs two barriers:
//  A. the group owning row k forms the norm of H[k][k+1:], alpha, v0 and
//     tau, and publishes v (zero at indices <= k) to LDS and to global
//     memory (barrier 1);
//  B. every thread multiplies its block by v; row sums give p = tau H v for
//     the group's rows, which is published with the group's share of p . v
//     (barrier 2);
//  C. every thread forms K = tau (p . v) / 2 and w = p - K v and applies
//     H -= v w^T + w v^T to its block in registers.
// v, p and the shares of p . v are double-buffered by the parity of k, so
// step k + 1 may overwrite nothing that a thread still reads from step k.
// Rows and columns <= k are finished: p is zero there, so the update leaves
// them unchanged, and register rows that lie wholly in them are skipped.
// The step's critical path is the owning group's reflector (a norm, a square
// root and a division in one wave, ~4,000 cycles), which the other waves
// wait for at barrier 1.
constexpr int TRI_NT = 512;

template <int CB>
__global__ void __launch_bounds__(TRI_NT)
k_rr_tri(int N, int nev, const double* __restrict__ G, int ldg, double* __restrict__ work,
         unsigned long long* __restrict__ prof) {
    constexpr int RA = CB, CC = 2 * CB;
    __shared__ double vbuf[2][NMAX], pbuf[2][NMAX], kpart[2][32];
    __shared__ double sd[NMAX], se[NMAX], stau[NMAX];
    RRWork W(work, N, nev);
    const int tid = threadIdx.x, g = tid >> 4, l = tid & 15;
    stamp(prof, 0);
    double h[RA][CC];
#pragma unroll
    for (int a = 0; a < RA; ++a) {
        const int i = g + 32 * a;
#pragma unroll
        for (int b = 0; b < CC; ++b) {
            const int j = l + 16 * b;
            h[a][b] = (i < N && j < N) ? 0.5 * (G[(int64_t)i * ldg + j] + G[(int64_t)j * ldg + i]) : 0.0;
        }
    }
    stamp(prof, 1);
    for (int k = 0; k < N - 2; ++k) {
        const int par = k & 1;
        // A. reflector of row k (owned by group k % 32, register row k / 32)
        if (g == (k & 31)) {
            const int ak = k >> 5;
            double x[CC];
#pragma unroll
            for (int b = 0; b < CC; ++b) x[b] = 0.0;
#pragma unroll
            for (int a = 0; a < RA; ++a)
                if (a == ak) {
#pragma unroll
                    for (int b = 0; b < CC; ++b) x[b] = h[a][b];
                }
            if (l == (k & 15)) {
#pragma unroll
                for (int b = 0; b < CC; ++b)
                    if (b == (k >> 4)) sd[k] = x[b];
            }
            double part = 0.0, x0l = 0.0;
#pragma unroll
            for (int b = 0; b < CC; ++b) {
                const int j = l + 16 * b;
                if (j <= k || j >= N) x[b] = 0.0;
                part += x[b] * x[b];
                if (b == ((k + 1) >> 4)) x0l = x[b];
            }
            const double xn2 = row_sum16(part);
            const int src = ((g & 3) << 4) + ((k + 1) & 15);  // lane of H[k][k+1] in this wave
            const double x0 = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x0l), src),
                                               __builtin_amdgcn_readlane(__double2loint(x0l), src));
            double alpha = 0.0, v0 = 0.0, t = 0.0;
            if (xn2 != 0.0) {
                // the short-latency forms on the step's critical path (the
                // exact ones for extreme magnitudes; xn2 is the same in every
                // lane, so one branch runs)
                if (xn2 > 1e-280 && xn2 < 1e280) {
                    const double sq = sqrt_nr(xn2);
                    alpha = (x0 >= 0.0) ? -sq : sq;
                    v0 = x0 - alpha;
                    t = 2.0 * rcp_nr(xn2 - x0 * x0 + v0 * v0);
                } else {
                    alpha = (x0 >= 0.0) ? -sqrt(xn2) : sqrt(xn2);
                    v0 = x0 - alpha;
                    t = 2.0 / (xn2 - x0 * x0 + v0 * v0);
                }
            }
#pragma unroll
            for (int b = 0; b < CC; ++b) {
                const int j = l + 16 * b;
                const double v = (t == 0.0) ? 0.0 : ((j == k + 1) ? v0 : x[b]);
                if (j < N) vbuf[par][j] = v;
                if (j > k && j < N) W.R[(int64_t)k * N + (j - k - 1)] = v;
            }
            if (l == 0) {
                se[k] = alpha;
                stau[k] = t;
            }
        }
        __syncthreads();
        const double t = stau[k];
        if (t == 0.0) continue;  // H untouched (uniform)
        // B. p = tau H v for this group's rows
        // register rows a < a_lo hold rows <= k only and are skipped; within
        // a row, finished columns get exact zero updates (v and w are zero
        // there), which costs less than a branch per block
        const int a_lo = (k >= 31) ? (k - 31) / 32 + 1 : 0;
        double vc[CC];
#pragma unroll
        for (int b = 0; b < CC; ++b) {
            const int j = l + 16 * b;
            vc[b] = (j < N) ? vbuf[par][j] : 0.0;
        }
        double kp = 0.0;
#pragma unroll
        for (int a = 0; a < RA; ++a) {
            const int i = g + 32 * a;
            double p = 0.0;
            if (a >= a_lo) {
                double s = 0.0;
#pragma unroll
                for (int b = 0; b < CC; ++b) s += h[a][b] * vc[b];
                p = t * row_sum16(s);
            }
            if (i <= k || i >= N) p = 0.0;
            if (i < N) kp += p * vbuf[par][i];
            if (l == a && i < N) pbuf[par][i] = p;
        }
        if (l == 0) kpart[par][g] = kp;
        __syncthreads();
        // C. rank-2 update H -= v w^T + w v^T, w = p - K v (v and p at this
        // thread's rows read back from LDS); every half-wave holds the 32
        // shares of p . v
        const double K = 0.5 * t * half_sum32(kpart[par][tid & 31]);
        double wc[CC];
#pragma unroll
        for (int b = 0; b < CC; ++b) {
            const int j = l + 16 * b;
            wc[b] = ((j < N) ? pbuf[par][j] : 0.0) - K * vc[b];
        }
#pragma unroll
        for (int a = 0; a < RA; ++a) {
            const int i = g + 32 * a;
            if (a < a_lo || i >= N) continue;
            const double vr = vbuf[par][i];
            const double wr = pbuf[par][i] - K * vr;
#pragma unroll
            for (int b = 0; b < CC; ++b)
                h[a][b] -= vr * wc[b] + wr * vc[b];
        }
    }
    // the last 2 x 2 block (or the 1 x 1 matrix)
#pragma unroll
    for (int a = 0; a < RA; ++a) {
        const int i = g + 32 * a;
#pragma unroll
        for (int b = 0; b < CC; ++b) {
            const int j = l + 16 * b;
            if (N >= 2 && i == N - 2 && j == N - 2) sd[N - 2] = h[a][b];
            if (N >= 2 && i == N - 1 && j == N - 2) se[N - 2] = h[a][b];
            if (i == N - 1 && j == N - 1) sd[N - 1] = h[a][b];
        }
    }
    if (tid == 0) {
        if (N >= 2) stau[N - 2] = 0.0;
        stau[N - 1] = 0.0;
        se[N - 1] = 0.0;
    }
    __syncthreads();
    for (int i = tid; i < N; i += TRI_NT) {
        W.d[i] = sd[i];
        W.e[i] = se[i];
        W.tau[i] = stau[i];
    }
    if (tid == 0) {
        double lo = sd[0], hi = sd[0], nrm = 0.0;
        for (int i = 0; i < N; ++i) {
            const double rr = (i > 0 ? fabs(se[i - 1]) : 0.0) + (i < N - 1 ? fabs(se[i]) : 0.0);
            lo = fmin(lo, sd[i] - rr);
            hi = fmax(hi, sd[i] + rr);
            nrm = fmax(nrm, fabs(sd[i]) + rr);
        }
        W.bnd[0] = lo;
        W.bnd[1] = hi;
        W.bnd[2] = fmax(nrm, 1e-300);
    }
    stamp(prof, 2);
}

// ---- 2. eigenvalues by multisection ----------------------------------------
__global__ void __launch_bounds__(64)
k_rr_bisect(int N, int nev, double* __restrict__ work, double* __restrict__ theta,
            unsigned long long* __restrict__ prof) {
    __shared__ double sd[NMAX], se2[NMAX];
    stamp(prof, 3);
    RRWork W(work, N, nev);
    const int j = blockIdx.x, q = threadIdx.x;
    // T scaled by a power of two (exact) to ||T|| <= 1
    const double sc = ldexp(1.0, -ilogb(W.bnd[2]) - 1);
    for (int i = q; i < N; i += 64) {
        sd[i] = W.d[i] * sc;
        se2[i] = (i < N - 1) ? (W.e[i] * sc) * (W.e[i] * sc) : 0.0;
    }
    __syncthreads();
    const double glo = W.bnd[0] * sc, ghi = W.bnd[1] * sc, tnorm = W.bnd[2] * sc;
    const double pivmin = 1e-290 + 1e-300 * tnorm;
    double a = glo - 1e-14 * tnorm, b = ghi + 1e-14 * tnorm;
    for (int it = 0; it < 24; ++it) {
        const double h = (b - a) / 65.0;
        const double x = a + (q + 1) * h;
        const bool pred = sturm(sd, se2, N, x, pivmin) > j;
        const unsigned long long bal = __ballot(pred);
        const int first = bal ? (__ffsll((long long)bal) - 1) : 64;
        const double nb = (first < 64) ? a + (first + 1) * h : b;
        const double na = (first > 0) ? a + first * h : a;
        a = na;
        b = nb;
        if (b - a <= 2.2e-16 * (fabs(a) + fabs(b)) + 1e-17 * tnorm) break;
    }
    if (q == 0) {
        const double lam = 0.5 * (a + b) / sc;
        W.lam[j] = lam;
        theta[j] = lam;
    }
}

// ---- 3. eigenvectors of T by inverse iteration ------------------------------
// lane q owns the entries i = q, q + 64, q + 128 of every vector it touches
__device__ __forceinline__ double wave_dot3(const double* y, const double* z, int N) {
    double s = 0.0;
    for (int i = threadIdx.x; i < N; i += 64) s += y[i] * z[i];
    return wsum(s);
}

__global__ void __launch_bounds__(64)
k_rr_invit(int N, int nev, double* __restrict__ work) {
    __shared__ double sd[NMAX], se[NMAX], ws[5 * NMAX], y[NMAX];
    RRWork W(work, N, nev);
    const int j = blockIdx.x, q = threadIdx.x;
    const double tnorm = W.bnd[2];
    const double* lam = W.lam;
    const double scale_j = fmax(fabs(lam[j]), 1e-6 * tnorm);
    const bool leader = (j == 0) || (lam[j] - lam[j - 1] > 1e-7 * scale_j);
    if (!leader) return;  // whole wave
    int gend = j + 1;
    while (gend < nev && lam[gend] - lam[gend - 1] <= 1e-7 * fmax(fabs(lam[gend]), 1e-6 * tnorm)) ++gend;
    // T and the shifts scaled by a power of two (exact) to ||T|| <= 1: the
    // iterates grow like 1 / (eps ||T||) and their squared norms must stay
    // finite and normal whatever the magnitude of H
    const double sc = ldexp(1.0, -ilogb(tnorm) - 1);
    for (int i = q; i < N; i += 64) {
        sd[i] = W.d[i] * sc;
        se[i] = W.e[i] * sc;
    }
    const double tiny = 2.2e-16 * tnorm * sc;
    for (int g = j; g < gend; ++g) {
        for (int i = q; i < N; i += 64) y[i] = 1.0 + 0.1 * (double)((7 * i + 3 * g) % 11);
        __syncthreads();
        if (q == 0) tri_factor(sd, se, N, lam[g] * sc + (double)(g - j) * 8.0 * tiny, tiny, ws);
        for (int itr = 0; itr < 2; ++itr) {
            __syncthreads();
            if (q == 0) tri_lu_solve(ws, N, y);
            __syncthreads();
            // Gram-Schmidt against the group's earlier vectors (Y columns
            // written by this same lane mapping), then normalise
            for (int h = j; h < g; ++h) {
                const double* yh = W.Y + (int64_t)h * N;
                const double s = wave_dot3(y, yh, N);
                for (int i = q; i < N; i += 64) y[i] -= s * yh[i];
            }
            const double nn = 1.0 / sqrt(fmax(wave_dot3(y, y, N), 1e-300));
            for (int i = q; i < N; i += 64) y[i] *= nn;
        }
        for (int i = q; i < N; i += 64) W.Y[(int64_t)g * N + i] = y[i];
    }
}

// ---- 4. back-transformation Z = Q Y ----------------------------------------
// One wave per eigenvector (rows q, q + 64, q + 128 in lane q), one
// workgroup each, so the vectors spread over CUs.  The N - 2 reflector steps
// are dependent; each one's entries are loaded BACK_PD steps ahead into a
// register ring, which hides the L2 latency, and the wave sums are DPP sums.
constexpr int BACK_PD = 4;

__global__ void __launch_bounds__(64)
k_rr_back(int N, int nev, double* __restrict__ work, unsigned long long* __restrict__ prof) {
    __shared__ double tau[NMAX];
    stamp(prof, 4);
    RRWork W(work, N, nev);
    const int g = blockIdx.x, q = threadIdx.x;
    for (int i = q; i < N; i += 64) tau[i] = W.tau[i];
    const double* yg = W.Y + (int64_t)g * N;
    double z0 = (q < N) ? yg[q] : 0.0;
    double z1 = (q + 64 < N) ? yg[q + 64] : 0.0;
    double z2 = (q + 128 < N) ? yg[q + 128] : 0.0;
    // reflector k acts on rows k+1 .. N-1 with v_k[row - k - 1] = R[k][row - k - 1]
    auto vload = [&](int k, int row) -> double {
        return (k >= 0 && row >= k + 1 && row < N) ? W.R[(int64_t)k * N + (row - k - 1)] : 0.0;
    };
    double ring[BACK_PD][3];
#pragma unroll
    for (int d = 0; d < BACK_PD; ++d) {
        const int k = N - 3 - d;
        ring[d][0] = vload(k, q);
        ring[d][1] = vload(k, q + 64);
        ring[d][2] = vload(k, q + 128);
    }
    __syncthreads();  // tau staged
    for (int k0 = N - 3; k0 >= 0; k0 -= BACK_PD) {
#pragma unroll
        for (int d = 0; d < BACK_PD; ++d) {
            const int k = k0 - d;
            if (k < 0) break;
            const double v0 = ring[d][0], v1 = ring[d][1], v2 = ring[d][2];
            ring[d][0] = vload(k - BACK_PD, q);
            ring[d][1] = vload(k - BACK_PD, q + 64);
            ring[d][2] = vload(k - BACK_PD, q + 128);
            const double t = tau[k];
            if (t == 0.0) continue;
            const double sdot = t * wsum(v0 * z0 + v1 * z1 + v2 * z2);
            z0 -= sdot * v0;
            z1 -= sdot * v1;
            z2 -= sdot * v2;
        }
    }
    if (q < N) W.Z[(int64_t)q * nev + g] = z0;
    if (q + 64 < N) W.Z[(int64_t)(q + 64) * nev + g] = z1;
    if (q + 128 < N) W.Z[(int64_t)(q + 128) * nev + g] = z2;
}

// ---- 5. orthonormalisation in the small space and output --------------------
// 4 x 4 tiles of Z in the update: with N nev + nev^2 <= PK and N <= NMAX
// there are at most 864 (ceil(N/4) ceil(nev/4), maximised over the
// admissible shapes), one per thread of the 1024
constexpr int POL_UT = 1;
__global__ void __launch_bounds__(NB)
k_rr_polish(int N, int nev, double* __restrict__ work, double* __restrict__ C, int* __restrict__ status,
            unsigned long long* __restrict__ prof) {
    __shared__ double Zs[PK];  // Z (N x nev), then E (nev x nev) behind it
    __shared__ double red[32];
    stamp(prof, 5);
    RRWork W(work, N, nev);
    const int tid = threadIdx.x;
    const int nz = N * nev;
    if (nz + nev * nev <= PK && ((N + 3) / 4) * ((nev + 3) / 4) <= NB * POL_UT) {
        double* E = Zs + nz;
        for (int idx = tid; idx < nz; idx += NB) Zs[idx] = W.Z[idx];
        __syncthreads();
        if (nev >= 32) {
            // both products in 4 x 4 register tiles: every LDS value loaded serves
            // four multiply-adds
            const int nt = (nev + 3) / 4, nr = (N + 3) / 4;
            const int ntri = nt * (nt + 1) / 2, nupd = nr * nt;
            for (int pass = 0; pass < 3; ++pass) {
                // E = Z^T Z - I: lower-triangle tiles, mirrored
                double emax = 0.0;
                for (int t = tid; t < ntri; t += NB) {
                    int ta = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
                    while ((ta + 1) * (ta + 2) / 2 <= t) ++ta;
                    while (ta * (ta + 1) / 2 > t) --ta;
                    const int a0 = 4 * ta, b0 = 4 * (t - ta * (ta + 1) / 2);
                    double sacc[4][4];
    #pragma unroll
                    for (int u = 0; u < 4; ++u)
    #pragma unroll
                        for (int w = 0; w < 4; ++w) sacc[u][w] = 0.0;
                    for (int i = 0; i < N; ++i) {
                        const double* zr = Zs + i * nev;
                        double za[4], zb[4];
    #pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            za[u] = (a0 + u < nev) ? zr[a0 + u] : 0.0;
                            zb[u] = (b0 + u < nev) ? zr[b0 + u] : 0.0;
                        }
    #pragma unroll
                        for (int u = 0; u < 4; ++u)
    #pragma unroll
                            for (int w = 0; w < 4; ++w) sacc[u][w] += za[u] * zb[w];
                    }
    #pragma unroll
                    for (int u = 0; u < 4; ++u)
    #pragma unroll
                        for (int w = 0; w < 4; ++w) {
                            const int a = a0 + u, b = b0 + w;
                            if (a >= nev || b > a) continue;
                            const double ev = sacc[u][w] - (a == b ? 1.0 : 0.0);
                            E[a * nev + b] = ev;
                            E[b * nev + a] = ev;
                            emax = fmax(emax, fabs(ev));
                        }
                }
                // block max of |E|
                for (int off = 32; off > 0; off >>= 1) emax = fmax(emax, __shfl_xor(emax, off, 64));
                __syncthreads();
                if ((tid & 63) == 0) red[tid >> 6] = emax;
                __syncthreads();
                double em = 0.0;
    #pragma unroll
                for (int w = 0; w < NB / 64; ++w) em = fmax(em, red[w]);
                // Newton-Schulz needs ||E|| < 1; a basis that far off is left as it is
                if (em >= 0.5 || (pass > 0 && em <= 1e-8)) break;
                // Z <- Z - Z E / 2: one 4 x 4 tile of Z per thread, written after a
                // barrier once every tile has been read
                const bool has = tid < nupd;
                const int i0 = 4 * (tid / nt), c0 = 4 * (tid % nt);
                double sacc[4][4];
    #pragma unroll
                for (int u = 0; u < 4; ++u)
    #pragma unroll
                    for (int w = 0; w < 4; ++w) sacc[u][w] = 0.0;
                if (has) {
                    for (int h = 0; h < nev; ++h) {
                        const double* er = E + h * nev;
                        double zi[4], eh[4];
    #pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            zi[u] = (i0 + u < N) ? Zs[(i0 + u) * nev + h] : 0.0;
                            eh[u] = (c0 + u < nev) ? er[c0 + u] : 0.0;
                        }
    #pragma unroll
                        for (int u = 0; u < 4; ++u)
    #pragma unroll
                            for (int w = 0; w < 4; ++w) sacc[u][w] += zi[u] * eh[w];
                    }
    #pragma unroll
                    for (int u = 0; u < 4; ++u)
    #pragma unroll
                        for (int w = 0; w < 4; ++w) {
                            const int i = i0 + u, c = c0 + w;
                            sacc[u][w] = (i < N && c < nev) ? Zs[i * nev + c] - 0.5 * sacc[u][w] : 0.0;
                        }
                }
                __syncthreads();
                if (has) {
    #pragma unroll
                    for (int u = 0; u < 4; ++u)
    #pragma unroll
                        for (int w = 0; w < 4; ++w) {
                            const int i = i0 + u, c = c0 + w;
                            if (i < N && c < nev) Zs[i * nev + c] = sacc[u][w];
                        }
                }
                __syncthreads();
                if (em <= 1e-8) break;  // one step leaves ~ em^2: rounding level
            }
        } else {
            // few columns: one output per thread keeps more threads busy
            for (int pass = 0; pass < 3; ++pass) {
                // E = Z^T Z - I (symmetric: lower entries computed, mirrored)
                double emax = 0.0;
                for (int idx = tid; idx < nev * nev; idx += NB) {
                    const int a = idx / nev, b = idx - a * nev;
                    if (b > a) continue;
                    double s0 = 0.0, s1 = 0.0;
                    int i = 0;
                    for (; i + 1 < N; i += 2) {
                        s0 += Zs[i * nev + a] * Zs[i * nev + b];
                        s1 += Zs[(i + 1) * nev + a] * Zs[(i + 1) * nev + b];
                    }
                    if (i < N) s0 += Zs[i * nev + a] * Zs[i * nev + b];
                    const double ev = (s0 + s1) - (a == b ? 1.0 : 0.0);
                    E[a * nev + b] = ev;
                    E[b * nev + a] = ev;
                    emax = fmax(emax, fabs(ev));
                }
                // block max of |E|
                for (int off = 32; off > 0; off >>= 1) emax = fmax(emax, __shfl_xor(emax, off, 64));
                __syncthreads();
                if ((tid & 63) == 0) red[tid >> 6] = emax;
                __syncthreads();
                double em = 0.0;
    #pragma unroll
                for (int w = 0; w < NB / 64; ++w) em = fmax(em, red[w]);
                // Newton-Schulz needs ||E|| < 1; a basis that far off is left as it is
                if (em >= 0.5 || (pass > 0 && em <= 1e-8)) break;
                // Z <- Z - Z E / 2 ; each thread owns entries idx = tid + NB * u
                constexpr int U = (PK + NB - 1) / NB;
                double nz_[U];
    #pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int idx = tid + NB * u;
                    nz_[u] = 0.0;
                    if (idx < nz) {
                        const int i = idx / nev, c = idx - i * nev;
                        double s = 0.0;
                        for (int h = 0; h < nev; ++h) s += Zs[i * nev + h] * E[h * nev + c];
                        nz_[u] = Zs[idx] - 0.5 * s;
                    }
                }
                __syncthreads();
    #pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int idx = tid + NB * u;
                    if (idx < nz) Zs[idx] = nz_[u];
                }
                __syncthreads();
                if (em <= 1e-8) break;  // one step leaves ~ em^2: rounding level
            }
        }
        for (int idx = tid; idx < nz; idx += NB) C[idx] = Zs[idx];
    } else {
        for (int idx = tid; idx < nz; idx += NB) C[idx] = W.Z[idx];
    }
    if (tid == 0) status[0] = 0;
    stamp(prof, 6);
}

}  // namespace

size_t rr_std_work_doubles(int N, int nev) { return (size_t)N * N + 3 * (size_t)N + 4 + nev + 2 * (size_t)N * nev + 64; }

void rr_std(lh_ctx* ctx, int N, int nev, const double* G, int ldg, double* C, double* theta, double* work,
            int* status, unsigned long long* prof) {
    if (N > kRRStdMaxN || nev > N || N < 1 || nev < 1)
        LH_THROW(LH_ERR_ARG, "rr_std: N=%d nev=%d out of range", N, nev);
    switch ((N + 31) / 32) {
        case 1: k_rr_tri<1><<<1, TRI_NT, 0, ctx->stream>>>(N, nev, G, ldg, work, prof); break;
        case 2: k_rr_tri<2><<<1, TRI_NT, 0, ctx->stream>>>(N, nev, G, ldg, work, prof); break;
        case 3: k_rr_tri<3><<<1, TRI_NT, 0, ctx->stream>>>(N, nev, G, ldg, work, prof); break;
        case 4: k_rr_tri<4><<<1, TRI_NT, 0, ctx->stream>>>(N, nev, G, ldg, work, prof); break;
        case 5: k_rr_tri<5><<<1, TRI_NT, 0, ctx->stream>>>(N, nev, G, ldg, work, prof); break;
        default: k_rr_tri<6><<<1, TRI_NT, 0, ctx->stream>>>(N, nev, G, ldg, work, prof); break;
    }
    LH_CHECK_LAUNCH();
    k_rr_bisect<<<nev, 64, 0, ctx->stream>>>(N, nev, work, theta, prof);
    LH_CHECK_LAUNCH();
    k_rr_invit<<<nev, 64, 0, ctx->stream>>>(N, nev, work);
    LH_CHECK_LAUNCH();
    k_rr_back<<<nev, 64, 0, ctx->stream>>>(N, nev, work, prof);
    LH_CHECK_LAUNCH();
    k_rr_polish<<<1, NB, 0, ctx->stream>>>(N, nev, work, C, status, prof);
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
        // stamps: tri start, after load, tri end, bisect start, back start,
        // polish start, polish end (kernel starts on the in-order stream)
        unsigned long long p[8];
        d2h(ctx, p, prof.p, sizeof(p));
        int khz = 100000;
        (void)hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device);
        for (int i = 0; i < 6; ++i) phase_us[i] = (double)(p[i + 1] - p[i]) * 1000.0 / (double)khz;
        phase_us[6] = 0.0;
    }
    HIPX(hipStreamSynchronize(ctx->stream));
    LH_API_END(ctx)
}
