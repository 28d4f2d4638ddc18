// Preconditioned conjugate gradients for (alpha*B + beta*A) x = b.
//
// Replaces the SuperLU solves of heat.diffusion (lapy/heat.py:204-227:
// (M_lumped + t A) u = b0) and Solver.poisson (lapy/solver.py:843-877:
// A_ff x_f = b_f after removing Dirichlet rows/columns).  The right-hand
// sides are solved together as a row-major block (one CG recurrence per
// column).  The operator K is formed once on the device on A's pattern
// (B diagonal or on the same pattern), with the Dirichlet rows and columns
// zeroed in it, so the iteration never reads the mask: a zero row and a zero
// Jacobi weight keep r, z, p and q at zero there.  With alpha == 0 and no
// Dirichlet rows (pure Neumann), A is singular with the constants as null
// space when its row sums vanish to rounding: the right-hand side is then
// projected onto the complement of the constants, each iteration removes the
// residual's mean before updating it (so rounding cannot build up a
// component CG cannot reduce), and the mean-free solution is returned
// (SuperLU returns an arbitrary member of the solution set).  Element matrices
// computed from float32 coordinates have row sums of ~1e-8 of the diagonal;
// such an A is regular (its smallest eigenvalue ~1e-9) and is solved as it
// stands, as SuperLU does, with no projection.
//
// One iteration is two launches (Jacobi or no preconditioner), and no kernel
// waits for another's blocks: each kernel reads the block partials the
// previous one wrote, every block summing them itself in the same fixed order
// (so all blocks agree on alpha and beta, and results do not depend on
// scheduling):
//   k_cg_dir(_ell): beta from the partials of r.z and r.r, then
//                q = K z + beta q, p = z + beta p (so q = K p without a
//                second product), and block partials of p.q (and of sum(r)
//                for a singular system);
//   k_cg_update: alpha from the partials of p.q, then x += alpha p,
//                r -= alpha q (its mean removed for a singular system),
//                z = D^-1 r, and block partials of r.z and r.r; block 0
//                counts the iteration and sets a done flag when every column
//                has converged.
// The partials of r.z and r.r alternate between two buffers by iteration
// parity, so the direction kernel never reads what the update of the same
// iteration writes.  With the Chebyshev-Jacobi preconditioner (a polynomial
// degree in the options) the update leaves z alone, deg - 1 fused Chebyshev
// steps on the ELL copy form z = p(D^-1 K) D^-1 r, and one more launch
// writes the partials of r.z.  K z is a product on the operator's ELL copy,
// one thread per row (CSR with LPR lanes per row when rows are long).  Systems of 4,096
// rows or more are solved in the locality order of the pattern
// (built once per pattern and kept with it), and each workgroup walks one
// contiguous range of rows, with the ranges of one XCD adjacent: the rows
// gathered by a range then sit in that XCD's L2.  The host checks the done
// flag once per chunk of iterations; iterations enqueued after convergence
// exit at once.
#include "lh_dense.hpp"

#include <chrono>
#include <cmath>

using namespace lh;

namespace {

constexpr int kBS = 256;    // threads per block of the set-up kernels
constexpr int kIBS = 1024;  // threads per block of the iteration kernels
constexpr int kMC = 8;      // right-hand sides per launch (wider blocks: several launches)

// K = beta * A + alpha * B on A's pattern
template <typename VT>
__global__ void k_form_k(int64_t nnz, const VT* __restrict__ av, const VT* __restrict__ bv, double alpha,
                         double beta, double* __restrict__ kv) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nnz) return;
    kv[s] = beta * (double)av[s] + (bv ? alpha * (double)bv[s] : 0.0);
}
template <typename VT>
__global__ void k_add_diag(int64_t n, const int32_t* __restrict__ diagpos, const VT* __restrict__ d, double alpha,
                           double* __restrict__ kv, int* __restrict__ missing) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (d[i] == VT(0)) return;
    if (diagpos[i] < 0) { atomicOr(missing, 1); return; }
    kv[diagpos[i]] += alpha * (double)d[i];
}
// Dirichlet rows and columns of K set to zero (row i: all of it; else the
// entries in Dirichlet columns)
__global__ void k_mask_k(int64_t n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cols,
                         const uint8_t* __restrict__ dm, double* __restrict__ kv) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool drow = dm[i] != 0;
    for (int32_t s = rowptr[i]; s < rowptr[i + 1]; ++s)
        if (drow || dm[cols[s]]) kv[s] = 0.0;
}
// Y = mask ? 0 : X (rows of a row-major n x m block)
__global__ void k_mask_rows(int64_t n, int m, const uint8_t* __restrict__ dm, const double* __restrict__ X,
                            double* __restrict__ Y) {
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n * m) return;
    Y[g] = (dm && dm[g / m]) ? 0.0 : X[g];
}
// no preconditioner: weight 1, except 0 where the Jacobi weight is 0
// (Dirichlet rows, whose row of K is zero)
__global__ void k_dinv_unit(int64_t n, double* __restrict__ dinv) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dinv[i] = dinv[i] != 0.0 ? 1.0 : 0.0;
}

// ---- block reductions -------------------------------------------------------
// Blocks exchange their partial sums through global memory between kernels
// (each kernel's partials are read by the next kernel), never inside one
// kernel: a device-scope fence per block (the usual last-block pattern) makes
// every block write back its L2 and costs more than the work on this GPU.

// sum of v over the block's threads in a fixed order (wave shuffles, then the
// waves in order); every thread gets the result
__device__ __forceinline__ double block_allsum(double v, double* sh /* blockDim / 64 */) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    __syncthreads();  // sh free (an earlier call's readers are done)
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += sh[k];
    return t;
}
// sum over nparts partials part[b * m + c] (b = 0 .. nparts-1), the same
// fixed order in every block; every thread gets the result
__device__ __forceinline__ double allsum_partials(const double* __restrict__ part, int nparts, int m, int c,
                                                  double* sh) {
    double t = 0.0;
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) t += part[(int64_t)b * m + c];
    return block_allsum(t, sh);
}

// two block sums with one pair of barriers (each formed as block_allsum forms
// it); sh holds 2 * blockDim / 64 doubles
__device__ __forceinline__ void block_allsum2(double a, double b, double* sh, double& ra, double& rb) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_down(a, o, 64);
        b += __shfl_down(b, o, 64);
    }
    const int nw = (int)(blockDim.x >> 6);
    __syncthreads();  // sh free (an earlier call's readers are done)
    if ((threadIdx.x & 63) == 0) {
        sh[threadIdx.x >> 6] = a;
        sh[nw + (threadIdx.x >> 6)] = b;
    }
    __syncthreads();
    double ta = 0.0, tb = 0.0;
    for (int k = 0; k < nw; ++k) {
        ta += sh[k];
        tb += sh[nw + k];
    }
    ra = ta;
    rb = tb;
}

// K such sums at once: every partial load is issued before the first
// reduction waits, and one pair of barriers serves all K (each sum is formed
// in the order allsum_partials uses, so the results are the same).  sh holds
// K * blockDim / 64 doubles.
template <int K>
__device__ __forceinline__ void allsum_partials_k(const double* const (&part)[K], const int (&np)[K], int m, int c,
                                                  double* sh, double (&out)[K]) {
    double t[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
        t[q] = 0.0;
        for (int b = threadIdx.x; b < np[q]; b += blockDim.x) t[q] += part[q][(int64_t)b * m + c];
    }
#pragma unroll
    for (int q = 0; q < K; ++q)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) t[q] += __shfl_down(t[q], o, 64);
    const int nw = (int)(blockDim.x >> 6);
    __syncthreads();  // sh free (an earlier call's readers are done)
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int q = 0; q < K; ++q) sh[q * nw + (threadIdx.x >> 6)] = t[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < K; ++q) {
        double r = 0.0;
        for (int k = 0; k < nw; ++k) r += sh[q * nw + k];
        out[q] = r;
    }
}

// CG state on the device (m columns).  Iteration k reads r_k.z_k and r_k.r_k
// from the partial buffers of parity k % 2 and the update writes those of
// r_{k+1} into parity (k + 1) % 2, so a kernel never reads partials that
// blocks of the same kernel are writing.
struct CgState {
    const double* tol2;  // rtol^2 |b|^2 per column
    double* rz[2];       // partials of r.z (nb_upd x m)
    double* rr[2];       // partials of r.r (nb_upd x m)
    double* pq;          // partials of p.q (nb_dir x m)
    double* sr;          // partials of sum(r) (nb_dir x m, singular systems)
    int nb_upd, nb_dir;
    int* status;         // [0]: done, [1]: iterations done (written by block 0)
    double* rr_out;      // r.r of the latest iterate (written by block 0)
};

// ---- direction: q = K z + beta q, p = z + beta p, partials of p.q ----------
// gridDim.x is a multiple of 8; logical block xcd_block() walks rows
// [lb * rpb, (lb + 1) * rpb)
template <int LPR, int MC, bool SING>
__global__ void __launch_bounds__(kIBS, (MC == 1 ? 2 : 1))
k_cg_dir(int64_t n, int m, int c0, int mc, int par, const int32_t* __restrict__ rowptr,
         const int32_t* __restrict__ cols, const double* __restrict__ kv, const double* __restrict__ Z,
         double* __restrict__ P, double* __restrict__ Q, const double* __restrict__ R, CgState s) {
    if (s.status[0]) return;  // converged: the enqueued iterations end here
    __shared__ double sh[4 * kIBS / 64];
    constexpr int GPB = kIBS / LPR;
    const int lane = threadIdx.x % LPR, grp = threadIdx.x / LPR;
    const int64_t rpb = (n + gridDim.x - 1) / gridDim.x;
    const int64_t rb = xcd_block() * rpb, re = rb + rpb < n ? rb + rpb : n;
    double bet[MC], pq[MC], sr[MC];
#pragma unroll
    for (int c = 0; c < MC; ++c) {
        bet[c] = 0.0;
        pq[c] = sr[c] = 0.0;
        if (c < mc) {
            // beta_k = rz_k / rz_{k-1}; 0 for a converged column (and at k = 0,
            // where P and Q hold zeros anyway)
            const double* const pp[3] = {s.rz[par], s.rz[par ^ 1], s.rr[par]};
            const int np[3] = {s.nb_upd, s.nb_upd, s.nb_upd};
            double v[3];
            allsum_partials_k<3>(pp, np, m, c0 + c, sh, v);
            const double rzk = v[0], rzo = v[1], rrk = v[2];
            bet[c] = (rrk > s.tol2[c0 + c] && rzo != 0.0) ? rzk / rzo : 0.0;
        }
    }
    for (int64_t i = rb + grp; i < re; i += GPB) {
        const int32_t r0 = rowptr[i], r1 = rowptr[i + 1];
        double acc[MC];
#pragma unroll
        for (int c = 0; c < MC; ++c) acc[c] = 0.0;
        for (int32_t q = r0 + lane; q < r1; q += LPR) {
            const int64_t j = cols[q];
            const double w = kv[q];
            const double* zr = Z + j * m + c0;
#pragma unroll
            for (int c = 0; c < MC; ++c)
                if (c < mc) acc[c] += w * zr[c];
        }
#pragma unroll
        for (int c = 0; c < MC; ++c)
#pragma unroll
            for (int o = LPR / 2; o > 0; o >>= 1) acc[c] += __shfl_xor(acc[c], o, LPR);
        if (lane == 0) {
            const int64_t o = i * m + c0;
#pragma unroll
            for (int c = 0; c < MC; ++c) {
                if (c >= mc) continue;
                const double pn = Z[o + c] + bet[c] * P[o + c];
                const double qn = acc[c] + bet[c] * Q[o + c];
                P[o + c] = pn;
                Q[o + c] = qn;
                pq[c] += pn * qn;
                if (SING) sr[c] += R[o + c];
            }
        }
    }
#pragma unroll
    for (int c = 0; c < MC; ++c) {
        if (c >= mc) continue;
        double bp, bs;
        block_allsum2(pq[c], SING ? sr[c] : 0.0, sh, bp, bs);
        if (threadIdx.x == 0) {
            s.pq[(int64_t)blockIdx.x * m + c0 + c] = bp;
            if (SING) s.sr[(int64_t)blockIdx.x * m + c0 + c] = bs;
        }
    }
}

// the same on the operator's ELL copy (W slots per row, padding slots point
// at the row itself with weight 0): one thread per row, the row's own values
// and its slots loaded before the gathers, which then go out together;
// 256-thread blocks, 8 waves per SIMD for W = 8.  A column's arithmetic (the
// slots summed in order, the same blocks and partials) does not depend on how
// many columns share the launch, so a block of right-hand sides gives what
// one solve per column gives.
constexpr int kEBS = 256;
template <int W, int MC, bool SING>
__global__ void __launch_bounds__(kEBS) __attribute__((amdgpu_waves_per_eu((W <= 8 && MC == 1) ? 8 : 4)))
k_cg_dir_ell(int64_t n, int m, int c0, int mc, int par, const int32_t* __restrict__ ecols,
             const double* __restrict__ ev, const double* __restrict__ Z, double* __restrict__ P,
             double* __restrict__ Q, const double* __restrict__ R, CgState s) {
    if (s.status[0]) return;
    __shared__ double sh[4 * kEBS / 64];
    const int64_t rpb = (n + gridDim.x - 1) / gridDim.x;
    const int64_t rb = xcd_block() * rpb, re = rb + rpb < n ? rb + rpb : n;
    double bet[MC], pq[MC], sr[MC];
#pragma unroll
    for (int c = 0; c < MC; ++c) {
        bet[c] = 0.0;
        pq[c] = sr[c] = 0.0;
        if (c < mc) {
            const double* const pp[3] = {s.rz[par], s.rz[par ^ 1], s.rr[par]};
            const int np[3] = {s.nb_upd, s.nb_upd, s.nb_upd};
            double v[3];
            allsum_partials_k<3>(pp, np, m, c0 + c, sh, v);
            const double rzk = v[0], rzo = v[1], rrk = v[2];
            bet[c] = (rrk > s.tol2[c0 + c] && rzo != 0.0) ? rzk / rzo : 0.0;
        }
    }
    for (int64_t i = rb + threadIdx.x; i < re; i += kEBS) {
        const int64_t o = i * m + c0;
        double zo[MC], po[MC], qo[MC], ro[MC], acc[MC];
#pragma unroll
        for (int c = 0; c < MC; ++c) {
            zo[c] = po[c] = qo[c] = ro[c] = acc[c] = 0.0;
            if (c < mc) {
                zo[c] = Z[o + c];
                po[c] = P[o + c];
                qo[c] = Q[o + c];
                if (SING) ro[c] = R[o + c];
            }
        }
        // slots in groups of G (all W at once for one column), in order
        constexpr int G = MC == 1 ? W : 8;
#pragma unroll
        for (int k0 = 0; k0 < W; k0 += G) {
            int32_t col[G];
            double val[G];
            const int4* cp = reinterpret_cast<const int4*>(ecols + i * W + k0);
            const double2* vp = reinterpret_cast<const double2*>(ev + i * W + k0);
#pragma unroll
            for (int k = 0; k < G / 4; ++k) {
                const int4 q4 = cp[k];
                col[4 * k] = q4.x;
                col[4 * k + 1] = q4.y;
                col[4 * k + 2] = q4.z;
                col[4 * k + 3] = q4.w;
            }
#pragma unroll
            for (int k = 0; k < G / 2; ++k) {
                const double2 v2 = vp[k];
                val[2 * k] = v2.x;
                val[2 * k + 1] = v2.y;
            }
            if (MC == 1) {
                double g[G];
#pragma unroll
                for (int k = 0; k < G; ++k) g[k] = Z[(int64_t)col[k] * m + c0];
#pragma unroll
                for (int k = 0; k < G; ++k) acc[0] += val[k] * g[k];
            } else {
#pragma unroll
                for (int k = 0; k < G; ++k) {
                    const double* zr = Z + (int64_t)col[k] * m + c0;
#pragma unroll
                    for (int c = 0; c < MC; ++c)
                        if (c < mc) acc[c] += val[k] * zr[c];
                }
            }
        }
#pragma unroll
        for (int c = 0; c < MC; ++c) {
            if (c >= mc) continue;
            const double pn = zo[c] + bet[c] * po[c];
            const double qn = acc[c] + bet[c] * qo[c];
            P[o + c] = pn;
            Q[o + c] = qn;
            pq[c] += pn * qn;
            if (SING) sr[c] += ro[c];
        }
    }
#pragma unroll
    for (int c = 0; c < MC; ++c) {
        if (c >= mc) continue;
        double bp, bs;
        block_allsum2(pq[c], SING ? sr[c] : 0.0, sh, bp, bs);
        if (threadIdx.x == 0) {
            s.pq[(int64_t)blockIdx.x * m + c0 + c] = bp;
            if (SING) s.sr[(int64_t)blockIdx.x * m + c0 + c] = bs;
        }
    }
}

// ---- update: x += alpha p, r = r - mean - alpha q, (z = D^-1 r) -------------
// alpha_k = rz_k / p.q for an active column (r_k.r_k > tol2), else 0
template <int MC, bool SING, bool JAC>
__global__ void __launch_bounds__(kIBS)
k_cg_update(int64_t n, int m, int c0, int mc, int par, const double* __restrict__ dinv, const double* __restrict__ P,
            const double* __restrict__ Q, double* __restrict__ X, double* __restrict__ R, double* __restrict__ Z,
            CgState s, int last_chunk) {
    if (s.status[0]) return;
    __shared__ double sh[4 * kIBS / 64];
    // contiguous rows per block, an even count; a thread takes two adjacent
    // rows per step (one 16-byte load per operand for one column)
    int64_t rpb = (n + gridDim.x - 1) / gridDim.x;
    rpb += rpb & 1;
    const int64_t rb = blockIdx.x * rpb, re = rb + rpb < n ? rb + rpb : n;
    double al[MC], mu[MC], rz[MC], rr[MC];
    int nact = 0;
#pragma unroll
    for (int c = 0; c < MC; ++c) {
        al[c] = mu[c] = rz[c] = rr[c] = 0.0;
        if (c < mc) {
            const int cc = c0 + c;
            const double* const pp[4] = {s.rz[par], s.rr[par], s.pq, s.sr};
            const int np[4] = {s.nb_upd, s.nb_upd, s.nb_dir, SING ? s.nb_dir : 0};
            double v[4];
            allsum_partials_k<4>(pp, np, m, cc, sh, v);
            const double rzk = v[0], rrk = v[1], pqk = v[2];
            const bool act = rrk > s.tol2[cc];
            nact += act ? 1 : 0;
            al[c] = (act && pqk != 0.0) ? rzk / pqk : 0.0;
            if (SING && act) mu[c] = v[3] / (double)n;
            if (blockIdx.x == 0 && threadIdx.x == 0) s.rr_out[cc] = rrk;
        }
    }
    for (int64_t i = rb + 2 * (int64_t)threadIdx.x; i < re; i += 2 * kIBS) {
        const bool two = i + 1 < re;
        double vr[2][MC], vq[2][MC], vp[2][MC], vx[2][MC], vd[2];
        if (MC == 1 && m == 1 && two) {
            const double2 r2 = *reinterpret_cast<const double2*>(R + i);
            const double2 q2 = *reinterpret_cast<const double2*>(Q + i);
            const double2 p2 = *reinterpret_cast<const double2*>(P + i);
            const double2 x2 = *reinterpret_cast<const double2*>(X + i);
            vr[0][0] = r2.x; vr[1][0] = r2.y;
            vq[0][0] = q2.x; vq[1][0] = q2.y;
            vp[0][0] = p2.x; vp[1][0] = p2.y;
            vx[0][0] = x2.x; vx[1][0] = x2.y;
            if (JAC) {
                const double2 d2 = *reinterpret_cast<const double2*>(dinv + i);
                vd[0] = d2.x;
                vd[1] = d2.y;
            } else {
                vd[0] = vd[1] = 0.0;
            }
        } else {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int64_t row = (h == 0 || two) ? i + h : i;  // an odd tail repeats row i, unused
                const int64_t o = row * m + c0;
                vd[h] = JAC ? dinv[row] : 0.0;
#pragma unroll
                for (int c = 0; c < MC; ++c) {
                    vr[h][c] = vq[h][c] = vp[h][c] = vx[h][c] = 0.0;
                    if (c < mc) {
                        vr[h][c] = R[o + c];
                        vq[h][c] = Q[o + c];
                        vp[h][c] = P[o + c];
                        vx[h][c] = X[o + c];
                    }
                }
            }
        }
        double nr[2][MC], nx[2][MC], nz[2][MC];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int c = 0; c < MC; ++c) {
                double r = vr[h][c];
                if (SING) r -= mu[c];
                r -= al[c] * vq[h][c];
                nr[h][c] = r;
                nx[h][c] = vx[h][c] + al[c] * vp[h][c];
                nz[h][c] = vd[h] * r;
                if (h == 0 || two) {
                    if (JAC) rz[c] += r * nz[h][c];
                    rr[c] += r * r;
                }
            }
        if (MC == 1 && m == 1 && two) {
            *reinterpret_cast<double2*>(R + i) = make_double2(nr[0][0], nr[1][0]);
            *reinterpret_cast<double2*>(X + i) = make_double2(nx[0][0], nx[1][0]);
            if (JAC) *reinterpret_cast<double2*>(Z + i) = make_double2(nz[0][0], nz[1][0]);
        } else {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                if (h == 1 && !two) continue;
                const int64_t o = (i + h) * m + c0;
#pragma unroll
                for (int c = 0; c < MC; ++c) {
                    if (c >= mc) continue;
                    R[o + c] = nr[h][c];
                    X[o + c] = nx[h][c];
                    if (JAC) Z[o + c] = nz[h][c];
                }
            }
        }
    }
#pragma unroll
    for (int c = 0; c < MC; ++c) {
        if (c >= mc) continue;
        double b1, b2;
        block_allsum2(rr[c], JAC ? rz[c] : 0.0, sh, b1, b2);
        if (threadIdx.x == 0) {
            s.rr[par ^ 1][(int64_t)blockIdx.x * m + c0 + c] = b1;
            if (JAC) s.rz[par ^ 1][(int64_t)blockIdx.x * m + c0 + c] = b2;
        }
    }
    // status for the host (and the early exit of later launches): every block
    // reached the same decision from the same partials; block 0 records it.
    // The columns of this launch are all converged only if nact == 0; the
    // iteration counts once, on its last column chunk.
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (c0 == 0) s.status[2] = 0;
        s.status[2] += nact;
        if (last_chunk) {
            if (s.status[2] == 0) s.status[0] = 1;
            else s.status[1] += 1;
        }
    }
}

// ---- partials of r.z (after an external preconditioner) into parity par ----
template <int MC>
__global__ void __launch_bounds__(kBS)
k_cg_rz(int64_t n, int m, int c0, int mc, const double* __restrict__ R, const double* __restrict__ Z,
        double* __restrict__ part, const int* __restrict__ status) {
    if (status && status[0]) return;
    __shared__ double sh[kBS / 64];
    double rz[MC];
#pragma unroll
    for (int c = 0; c < MC; ++c) rz[c] = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) {
        const int64_t o = i * m + c0;
#pragma unroll
        for (int c = 0; c < MC; ++c)
            if (c < mc) rz[c] += R[o + c] * Z[o + c];
    }
#pragma unroll
    for (int c = 0; c < MC; ++c) {
        if (c >= mc) continue;
        const double b = block_allsum(rz[c], sh);
        if (threadIdx.x == 0) part[(int64_t)blockIdx.x * m + c0 + c] = b;
    }
}

// ---- Chebyshev-Jacobi preconditioner: z = p(D^-1 K) D^-1 r -------------------
// The Chebyshev semi-iteration (two-step form) for K z = r on [amin, bmax],
// started from z_0 = 0: z_1 = a D^-1 r, then
//   z_{k+1} = w_k (z_k - z_{k-1} + a D^-1 (r - K z_k)) + z_{k-1},
// a = 2 / (bmax + amin), written over z_{k-1} (row-local), z_k gathered on
// the operator's ELL copy (padding slots point at the row with weight 0).
// The map r -> z is a fixed polynomial in D^-1 K times D^-1: symmetric, and
// positive definite when p > 0 on the spectrum, which the residual
// polynomial's bound below 1 on [amin, bmax] gives (and p(0) > 0 below it).
// Converged solves (status[0]) skip the work.
__global__ void __launch_bounds__(kBS)
k_pc_cheb_init(int64_t n, int m, const double* __restrict__ dinv, const double* __restrict__ R, double a,
               double* __restrict__ Z1, double* __restrict__ Z0, const int* __restrict__ status) {
    if (status && status[0]) return;
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n * m) return;
    const int64_t i = g / m;
    Z1[g] = a * dinv[i] * R[g];
    Z0[g] = 0.0;
}

template <int W>
__global__ void __launch_bounds__(kEBS)
k_pc_cheb_step(int64_t n, int m, const int32_t* __restrict__ ecols, const double* __restrict__ ev,
               const double* __restrict__ dinv, const double* __restrict__ R, const double* __restrict__ Zk,
               double* __restrict__ Zp, double a, double w, const int* __restrict__ status) {
    if (status && status[0]) return;
    const int64_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t js[W];
    double ws[W];
#pragma unroll
    for (int q = 0; q < W; ++q) {
        js[q] = ecols[i * W + q];
        ws[q] = ev[i * W + q];
    }
    const double ad = a * dinv[i];
    for (int c = 0; c < m; ++c) {
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < W; ++q) acc += ws[q] * Zk[(int64_t)js[q] * m + c];
        const int64_t o = i * m + c;
        const double zp = Zp[o];
        Zp[o] = w * (Zk[o] - zp + ad * (R[o] - acc)) + zp;
    }
}

// nb_dir block maxima of the Gershgorin bound sum_j |K_ij| dinv_i of D^-1 K
template <int LPR>
__global__ void __launch_bounds__(kBS)
k_gersh(int64_t n, const int32_t* __restrict__ rowptr, const double* __restrict__ kv,
        const double* __restrict__ dinv, double* __restrict__ part) {
    __shared__ double sh[kBS / 64];
    constexpr int GPB = kBS / LPR;
    const int lane = threadIdx.x % LPR;
    double mx = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * GPB + threadIdx.x / LPR; i < n; i += (int64_t)gridDim.x * GPB) {
        double acc = 0.0;
        for (int32_t q = rowptr[i] + lane; q < rowptr[i + 1]; q += LPR) acc += fabs(kv[q]);
#pragma unroll
        for (int o = LPR / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, LPR);
        if (lane == 0) mx = fmax(mx, acc * dinv[i]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_down(mx, o, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = 0.0;
        for (int k = 0; k < kBS / 64; ++k) b = fmax(b, sh[k]);
        part[blockIdx.x] = b;
    }
}

// ---- set-up reductions -------------------------------------------------------
// partials of sum_i A[i, c] * (B ? B[i, c] : 1) (one per block), and the
// total from them by a one-block kernel
template <int MC>
__global__ void __launch_bounds__(kBS)
k_coldot_part(int64_t n, int m, int c0, int mc, const double* __restrict__ A, const double* __restrict__ B,
              double* __restrict__ part) {
    __shared__ double sh[kBS / 64];
    double acc[MC];
#pragma unroll
    for (int c = 0; c < MC; ++c) acc[c] = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) {
        const int64_t o = i * m + c0;
#pragma unroll
        for (int c = 0; c < MC; ++c)
            if (c < mc) acc[c] += A[o + c] * (B ? B[o + c] : 1.0);
    }
#pragma unroll
    for (int c = 0; c < MC; ++c) {
        if (c >= mc) continue;
        const double b = block_allsum(acc[c], sh);
        if (threadIdx.x == 0) part[(int64_t)blockIdx.x * m + c0 + c] = b;
    }
}
__global__ void __launch_bounds__(kBS)
k_sum_parts(int nparts, int m, const double* __restrict__ part, double* __restrict__ out) {
    __shared__ double sh[kBS / 64];
    for (int c = 0; c < m; ++c) {
        const double t = allsum_partials(part, nparts, m, c, sh);
        if (threadIdx.x == 0) out[c] = t;
    }
}
// out[i, :] = in[perm[i], :] (row-major n x m; perm: new row -> old row)
__global__ void k_perm_rows(int64_t n, int m, const int32_t* __restrict__ perm, const double* __restrict__ in,
                            double* __restrict__ out) {
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n * m) return;
    const int64_t i = g / m;
    out[g] = in[(int64_t)perm[i] * m + (g - i * m)];
}
// X[:, c] -= sums[c] / n
__global__ void k_sub_colmean(int64_t n, int m, double* __restrict__ x, const double* __restrict__ sums) {
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n * m) return;
    x[g] -= sums[g % m] / (double)n;
}
// Y = B - K X (row-major n x m, LPR lanes per row)
template <int LPR>
__global__ void __launch_bounds__(kBS)
k_resid(int64_t n, int m, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cols,
        const double* __restrict__ kv, const double* __restrict__ X, const double* __restrict__ B,
        double* __restrict__ Y) {
    constexpr int GPB = kBS / LPR;
    const int lane = threadIdx.x % LPR;
    const int64_t i = (int64_t)blockIdx.x * GPB + threadIdx.x / LPR;
    if (i >= n) return;  // whole LPR group
    for (int c = 0; c < m; ++c) {
        double acc = 0.0;
        for (int32_t q = rowptr[i] + lane; q < rowptr[i + 1]; q += LPR) acc += kv[q] * X[(int64_t)cols[q] * m + c];
#pragma unroll
        for (int o = LPR / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, LPR);
        if (lane == 0) Y[i * m + c] = B[i * m + c] - acc;
    }
}
// per-block maxima of |(K 1)_i| * dinv_i (row sums relative to the diagonal)
template <int LPR>
__global__ void __launch_bounds__(kBS)
k_rowsum_rel(int64_t n, const int32_t* __restrict__ rowptr, const double* __restrict__ kv,
             const double* __restrict__ dinv, double* __restrict__ part) {
    __shared__ double sh[kBS / 64];
    constexpr int GPB = kBS / LPR;
    const int lane = threadIdx.x % LPR;
    double mx = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * GPB + threadIdx.x / LPR; i < n; i += (int64_t)gridDim.x * GPB) {
        double acc = 0.0;
        for (int32_t q = rowptr[i] + lane; q < rowptr[i + 1]; q += LPR) acc += kv[q];
#pragma unroll
        for (int o = LPR / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, LPR);
        if (lane == 0) mx = fmax(mx, fabs(acc * dinv[i]));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_down(mx, o, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = 0.0;
        for (int k = 0; k < kBS / 64; ++k) b = fmax(b, sh[k]);
        part[blockIdx.x] = b;
    }
}

double now_us() {
    return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// lanes per row for rows of about `avg` entries
int lanes_per_row(double avg) { return avg <= 4.0 ? 4 : avg <= 8.0 ? 8 : avg <= 16.0 ? 16 : 32; }

struct Launcher {
    lh_ctx* ctx;
    int64_t n;
    int m;
    int lpr;
    unsigned nb;      // blocks of the row-wise kernels (partials per column)
    unsigned nb_dir;  // blocks of k_cg_dir
    const int32_t* rp;
    const int32_t* cl;
    const double* kv;

    int ew = 0;  // ELL width of the operator's copy (0: CSR kernels)
    unsigned nb_ell = 0;  // blocks of k_cg_dir_ell
    const int32_t* ecols = nullptr;
    const double* ev = nullptr;

    template <int MC>
    void dir_mc(int c0, int mc, int par, bool sing, const double* Z, double* P, double* Q, const double* R,
                const CgState& s) {
        if constexpr (MC <= 4) {
          if (ew) {
#define LH_DIRE(W_, S) \
    k_cg_dir_ell<W_, MC, S><<<nb_ell, kEBS, 0, ctx->stream>>>(n, m, c0, mc, par, ecols, ev, Z, P, Q, R, s)
            switch (ew) {
                case 8: if (sing) LH_DIRE(8, true); else LH_DIRE(8, false); break;
                case 16: if (sing) LH_DIRE(16, true); else LH_DIRE(16, false); break;
                case 24: if (sing) LH_DIRE(24, true); else LH_DIRE(24, false); break;
                default: if (sing) LH_DIRE(32, true); else LH_DIRE(32, false); break;
            }
#undef LH_DIRE
            LH_CHECK_LAUNCH();
            return;
          }
        }
#define LH_DIR(L, S) \
    k_cg_dir<L, MC, S><<<nb_dir, kIBS, 0, ctx->stream>>>(n, m, c0, mc, par, rp, cl, kv, Z, P, Q, R, s)
        switch (lpr) {
            case 4: if (sing) LH_DIR(4, true); else LH_DIR(4, false); break;
            case 8: if (sing) LH_DIR(8, true); else LH_DIR(8, false); break;
            case 16: if (sing) LH_DIR(16, true); else LH_DIR(16, false); break;
            default: if (sing) LH_DIR(32, true); else LH_DIR(32, false); break;
        }
#undef LH_DIR
        LH_CHECK_LAUNCH();
    }
    void dir(int par, bool sing, const double* Z, double* P, double* Q, const double* R, const CgState& s) {
        const int w = ew ? 4 : kMC;  // the ELL kernel takes four columns at most (eight would spill)
        for (int c0 = 0; c0 < m; c0 += w) {
            const int mc = std::min(w, m - c0);
            if (mc == 1) dir_mc<1>(c0, mc, par, sing, Z, P, Q, R, s);
            else if (mc <= 4) dir_mc<4>(c0, mc, par, sing, Z, P, Q, R, s);
            else dir_mc<8>(c0, mc, par, sing, Z, P, Q, R, s);
        }
    }
    template <int MC>
    void update_mc(int c0, int mc, int par, bool sing, bool jac, const double* dinv, const double* P,
                   const double* Q, double* X, double* R, double* Z, const CgState& s, int last) {
#define LH_UPD(S, J) \
    k_cg_update<MC, S, J><<<nb, kIBS, 0, ctx->stream>>>(n, m, c0, mc, par, dinv, P, Q, X, R, Z, s, last)
        if (sing) { if (jac) LH_UPD(true, true); else LH_UPD(true, false); }
        else { if (jac) LH_UPD(false, true); else LH_UPD(false, false); }
#undef LH_UPD
        LH_CHECK_LAUNCH();
    }
    // two columns per launch at most (two rows of four columns would spill)
    void update(int par, bool sing, bool jac, const double* dinv, const double* P, const double* Q, double* X,
                double* R, double* Z, const CgState& s) {
        for (int c0 = 0; c0 < m; c0 += 2) {
            const int mc = std::min(2, m - c0);
            const int last = c0 + 2 >= m;
            if (mc == 1) update_mc<1>(c0, mc, par, sing, jac, dinv, P, Q, X, R, Z, s, last);
            else update_mc<2>(c0, mc, par, sing, jac, dinv, P, Q, X, R, Z, s, last);
        }
    }
    // partials (nb per column) of r.z into part
    void rz(const double* R, const double* Z, double* part, const int* status) {
        for (int c0 = 0; c0 < m; c0 += kMC) {
            const int mc = std::min(kMC, m - c0);
            if (mc == 1) k_cg_rz<1><<<nb, kBS, 0, ctx->stream>>>(n, m, c0, mc, R, Z, part, status);
            else if (mc <= 4) k_cg_rz<4><<<nb, kBS, 0, ctx->stream>>>(n, m, c0, mc, R, Z, part, status);
            else k_cg_rz<8><<<nb, kBS, 0, ctx->stream>>>(n, m, c0, mc, R, Z, part, status);
            LH_CHECK_LAUNCH();
        }
    }
    // partials (nb per column) of sum_i A[i, c] B[i, c] (B null: sum of A)
    void coldot_part(const double* A, const double* B, double* part) {
        for (int c0 = 0; c0 < m; c0 += kMC) {
            const int mc = std::min(kMC, m - c0);
            if (mc == 1) k_coldot_part<1><<<nb, kBS, 0, ctx->stream>>>(n, m, c0, mc, A, B, part);
            else if (mc <= 4) k_coldot_part<4><<<nb, kBS, 0, ctx->stream>>>(n, m, c0, mc, A, B, part);
            else k_coldot_part<8><<<nb, kBS, 0, ctx->stream>>>(n, m, c0, mc, A, B, part);
            LH_CHECK_LAUNCH();
        }
    }
    // out[c] = sum_i A[i, c] B[i, c] (device, m doubles)
    void coldot(const double* A, const double* B, double* part, double* out) {
        coldot_part(A, B, part);
        k_sum_parts<<<1, kBS, 0, ctx->stream>>>((int)nb, m, part, out);
        LH_CHECK_LAUNCH();
    }
    void resid(const double* X, const double* B, double* Y) {
        const unsigned g = nblocks(n * lpr, kBS);
        switch (lpr) {
            case 4: k_resid<4><<<g, kBS, 0, ctx->stream>>>(n, m, rp, cl, kv, X, B, Y); break;
            case 8: k_resid<8><<<g, kBS, 0, ctx->stream>>>(n, m, rp, cl, kv, X, B, Y); break;
            case 16: k_resid<16><<<g, kBS, 0, ctx->stream>>>(n, m, rp, cl, kv, X, B, Y); break;
            default: k_resid<32><<<g, kBS, 0, ctx->stream>>>(n, m, rp, cl, kv, X, B, Y); break;
        }
        LH_CHECK_LAUNCH();
    }
    // nb_dir block maxima of sum_j |K_ij| dinv_i
    void gersh(const double* dinv, double* part) {
        switch (lpr) {
            case 4: k_gersh<4><<<nb_dir, kBS, 0, ctx->stream>>>(n, rp, kv, dinv, part); break;
            case 8: k_gersh<8><<<nb_dir, kBS, 0, ctx->stream>>>(n, rp, kv, dinv, part); break;
            case 16: k_gersh<16><<<nb_dir, kBS, 0, ctx->stream>>>(n, rp, kv, dinv, part); break;
            default: k_gersh<32><<<nb_dir, kBS, 0, ctx->stream>>>(n, rp, kv, dinv, part); break;
        }
        LH_CHECK_LAUNCH();
    }
    // Z = p(D^-1 K) D^-1 R (degree deg, on [amin, bmax]) with T as the second
    // work block; needs the ELL copy
    void cheb(int deg, double amin, double bmax, const double* dinv, const double* R, double* Z, double* T,
              const int* status) {
        const double a = 2.0 / (bmax + amin), mu = (bmax - amin) / (bmax + amin);
        // the last iterate must land in Z: x_1 goes to Z when deg - 1 is even
        double* zk = ((deg - 1) % 2 == 0) ? Z : T;
        double* zp = (zk == Z) ? T : Z;
        k_pc_cheb_init<<<nblocks(n * m, kBS), kBS, 0, ctx->stream>>>(n, m, dinv, R, a, zk, zp, status);
        LH_CHECK_LAUNCH();
        double w = 1.0;
        for (int k = 1; k < deg; ++k) {
            w = (k == 1) ? 1.0 / (1.0 - 0.5 * mu * mu) : 1.0 / (1.0 - 0.25 * mu * mu * w);
        const unsigned nbr = nblocks_xcd(n, kEBS);  // one thread per row
#define LH_PCS(W_) \
    k_pc_cheb_step<W_><<<nbr, kEBS, 0, ctx->stream>>>(n, m, ecols, ev, dinv, R, zk, zp, a, w, status)
            switch (ew) {
                case 8: LH_PCS(8); break;
                case 16: LH_PCS(16); break;
                case 24: LH_PCS(24); break;
                default: LH_PCS(32); break;
            }
#undef LH_PCS
            LH_CHECK_LAUNCH();
            std::swap(zk, zp);
        }
    }
    // nb_dir block maxima of |(K 1)_i| dinv_i
    void rowsum_rel(const double* dinv, double* part) {
        switch (lpr) {
            case 4: k_rowsum_rel<4><<<nb_dir, kBS, 0, ctx->stream>>>(n, rp, kv, dinv, part); break;
            case 8: k_rowsum_rel<8><<<nb_dir, kBS, 0, ctx->stream>>>(n, rp, kv, dinv, part); break;
            case 16: k_rowsum_rel<16><<<nb_dir, kBS, 0, ctx->stream>>>(n, rp, kv, dinv, part); break;
            default: k_rowsum_rel<32><<<nb_dir, kBS, 0, ctx->stream>>>(n, rp, kv, dinv, part); break;
        }
        LH_CHECK_LAUNCH();
    }
};

}  // namespace

extern "C" {

int lh_pcg(lh_ctx* ctx, lh_csr* A, lh_csr* B, double alpha, double beta, const double* b, int nrhs,
           const uint8_t* dmask, double* x, const lh_pcg_opts* opts, lh_stats* stats) {
    if (!ctx || !A || !b || !x || nrhs <= 0) return LH_ERR_ARG;
    LH_API_BEGIN(ctx)
    HIPX(hipSetDevice(ctx->device));
    const double t0 = now_us();
    const int64_t n = A->n(), nnz = A->nnz();
    if (B && B->n() != n) LH_THROW(LH_ERR_ARG, "A and B differ in size");
    if (alpha != 0.0 && !B) LH_THROW(LH_ERR_ARG, "alpha != 0 needs B");
    const int m = nrhs;
    const double rtol = (opts && opts->rtol > 0) ? opts->rtol : 1e-10;
    const int64_t max_iter = (opts && opts->max_iter > 0) ? opts->max_iter : std::max<int64_t>(1000, 10 * n);
    const int verbose = opts ? opts->verbose : 0;
    const int precond = opts ? opts->precond : LH_PRECOND_JACOBI;
    if (precond != LH_PRECOND_JACOBI && precond != LH_PRECOND_LEARNED && precond != LH_PRECOND_NONE)
        LH_THROW(LH_ERR_ARG, "preconditioner %d unknown", precond);
    // Chebyshev-Jacobi degree (Jacobi only): 0 or 1 plain Jacobi
    const int pdeg_opt = opts ? opts->poly_degree : 0;
    if (pdeg_opt < 0 || pdeg_opt > 256) LH_THROW(LH_ERR_ARG, "poly_degree must be in [0, 256], got %d", pdeg_opt);
    if (precond == LH_PRECOND_LEARNED && !ctx->mpnn)
        LH_THROW(LH_ERR_ARG, "LH_PRECOND_LEARNED needs a network on the context (lh_ctx_set_mpnn)");
    if ((size_t)m * 16 + 64 > ctx->host_small_bytes) LH_THROW(LH_ERR_ARG, "too many right-hand sides (%d)", m);
    lh_stats st{};
    // operator values on A's pattern
    DevBuf kv(nnz * 8), missing(8);
    memset_d(ctx, missing.p, 0, 4);
    const bool use_b = (alpha != 0.0);
    if (use_b && !B->is_diag && B->pat != A->pat)
        LH_THROW(LH_ERR_ARG, "B must be diagonal or share A's sparsity pattern");
    if (A->dtype == LH_F64) {
        k_form_k<double><<<nblocks(nnz, 256), 256, 0, ctx->stream>>>(
            nnz, A->vals.as<double>(), (use_b && !B->is_diag) ? B->vals.as<double>() : nullptr, alpha, beta,
            kv.as<double>());
        if (use_b && B->is_diag)
            k_add_diag<double><<<nblocks(n, 256), 256, 0, ctx->stream>>>(n, A->pat->diagpos.as<int32_t>(),
                                                                        B->dense_diag.as<double>(), alpha,
                                                                        kv.as<double>(), missing.as<int>());
    } else {
        k_form_k<float><<<nblocks(nnz, 256), 256, 0, ctx->stream>>>(
            nnz, A->vals.as<float>(), (use_b && !B->is_diag) ? B->vals.as<float>() : nullptr, alpha, beta,
            kv.as<double>());
        if (use_b && B->is_diag)
            k_add_diag<float><<<nblocks(n, 256), 256, 0, ctx->stream>>>(n, A->pat->diagpos.as<int32_t>(),
                                                                       B->dense_diag.as<float>(), alpha,
                                                                       kv.as<double>(), missing.as<int>());
    }
    LH_CHECK_LAUNCH();
    int miss = 0;
    d2h(ctx, &miss, missing.p, 4);
    if (miss) LH_THROW(LH_ERR_ARG, "B has diagonal entries outside A's pattern");
    bool any_dirichlet = false;
    if (dmask)
        for (int64_t i = 0; i < n; ++i)
            if (dmask[i]) { any_dirichlet = true; break; }
    DevBuf dm;
    const uint8_t* dmp = nullptr;
    if (any_dirichlet) {
        dm.alloc(n);
        h2d(ctx, dm.p, dmask, n);
        dmp = dm.as<uint8_t>();
        k_mask_k<<<nblocks(n, 256), 256, 0, ctx->stream>>>(n, A->pat->rowptr.as<int32_t>(),
                                                           A->pat->cols.as<int32_t>(), dmp, kv.as<double>());
        LH_CHECK_LAUNCH();
    }
    // the locality order of the pattern (not with the learned network, which
    // runs on A's own order): K permuted once per solve, the order itself
    // built once per pattern
    std::shared_ptr<Reorder> ro;
    if (n >= 4096 && precond != LH_PRECOND_LEARNED) {
        ro = pattern_reorder(ctx, *A->pat);
    }
    std::unique_ptr<lh_csr> Kp;
    const Pattern* kpat = A->pat.get();
    const double* kvp = kv.as<double>();
    if (ro) {
        lh_csr K;  // view of the operator's values on A's pattern
        K.ctx = ctx;
        K.pat = A->pat;
        K.dtype = LH_F64;
        K.vals.p = kv.p;
        Kp = ro->apply(ctx, &K);
        K.vals.p = nullptr;  // not owned
        kpat = Kp->pat.get();
        kvp = Kp->vals.as<double>();
    }
    // Jacobi weights from the (masked) operator's diagonal: 0 on Dirichlet rows
    DevBuf dinv(n * 8);
    {
        lh_csr K;  // view for jacobi_diag
        K.ctx = ctx;
        K.pat = ro ? Kp->pat : A->pat;
        K.dtype = LH_F64;
        K.vals.p = const_cast<double*>(kvp);
        jacobi_diag(ctx, &K, nullptr, 0.0, 1.0, dinv.as<double>());
        K.vals.p = nullptr;  // not owned
    }
    Launcher L;
    L.ctx = ctx;
    L.n = n;
    L.m = m;
    L.lpr = lanes_per_row(n > 0 ? (double)nnz / (double)n : 1.0);
    L.nb = std::max(1u, std::min<unsigned>(nblocks(n, 2 * kIBS), (unsigned)ctx->n_cu));
    L.nb_dir = std::max<unsigned>(kXCDs, std::min<unsigned>(nblocks_xcd(n * L.lpr, kIBS), (unsigned)ctx->n_cu * 2));
    L.nb_dir = (L.nb_dir + kXCDs - 1) / kXCDs * kXCDs;
    L.rp = kpat->rowptr.as<int32_t>();
    L.cl = kpat->cols.as<int32_t>();
    L.kv = kvp;
    // ELL copy of K for the direction kernel (rows of at most 32 entries)
    Ell kell;
    {
        lh_csr K;  // view
        K.ctx = ctx;
        K.pat = ro ? Kp->pat : A->pat;
        K.dtype = LH_F64;
        K.vals.p = const_cast<double*>(kvp);
        kell.build(ctx, &K, nullptr);
        K.vals.p = nullptr;  // not owned
    }
    if (kell.W) {
        L.ew = kell.W;
        L.ecols = kell.cols.as<int32_t>();
        L.ev = kell.va.as<double>();
        L.nb_ell = std::max<unsigned>(kXCDs, std::min<unsigned>(nblocks_xcd(n, kEBS), (unsigned)ctx->n_cu * 8));
        L.nb_ell = (L.nb_ell + kXCDs - 1) / kXCDs * kXCDs;
        L.nb_dir = L.nb_ell;  // the p.q partials come from the ELL launch
    }
    // device state
    const size_t pu = (size_t)L.nb * m, pd = (size_t)L.nb_dir * m;
    DevBuf parts((5 * pu + 2 * pd) * 8), scal((size_t)3 * m * 8), ints(16);
    double* pp = parts.as<double>();
    CgState s;
    s.rz[0] = pp;
    s.rz[1] = pp + pu;
    s.rr[0] = pp + 2 * pu;
    s.rr[1] = pp + 3 * pu;
    double* ptmp = pp + 4 * pu;  // set-up reductions
    s.pq = pp + 5 * pu;
    s.sr = s.pq + pd;
    s.nb_upd = (int)L.nb;
    s.nb_dir = (int)L.nb_dir;
    double* tol2 = scal.as<double>();
    double* sums = tol2 + m;
    double* rrout = sums + m;
    s.tol2 = tol2;
    s.rr_out = rrout;
    s.status = ints.as<int>();
    memset_d(ctx, parts.p, 0, parts.bytes);
    memset_d(ctx, ints.p, 0, 16);
    // pure Neumann: singular (constants as null space) only if max_i |(K 1)_i| / K_ii is at rounding level
    bool singular = false;
    if (alpha == 0.0 && !any_dirichlet) {
        L.rowsum_rel(dinv.as<double>(), s.pq);
        std::vector<double> hmax(L.nb_dir);
        d2h(ctx, hmax.data(), s.pq, (size_t)L.nb_dir * 8);
        double worst = 0.0;
        for (double v : hmax) worst = std::max(worst, v);
        singular = worst <= 1e-12;
    }
    if (precond == LH_PRECOND_NONE)
        k_dinv_unit<<<nblocks(n, 256), 256, 0, ctx->stream>>>(n, dinv.as<double>());
    // Chebyshev-Jacobi: degree, and the interval [bmax / (2 deg^2), bmax] with
    // bmax the Gershgorin bound of D^-1 K (above its largest eigenvalue)
    const int pdeg = (precond == LH_PRECOND_JACOBI && L.ew && pdeg_opt > 1) ? pdeg_opt : 1;
    double cheb_amin = 0.0, cheb_bmax = 2.0;
    if (pdeg > 1) {
        L.gersh(dinv.as<double>(), s.pq);
        std::vector<double> hg(L.nb_dir);
        d2h(ctx, hg.data(), s.pq, (size_t)L.nb_dir * 8);
        double g = 0.0;
        for (double v : hg) g = std::max(g, v);
        cheb_bmax = g > 0.0 ? g : 2.0;
        cheb_amin = cheb_bmax / (2.0 * pdeg * pdeg);
    }
    DevBuf dT;
    if (pdeg > 1) dT.alloc((size_t)n * m * 8);
    bool learned_on = false;
    LearnedPrecond learned;
    if (precond == LH_PRECOND_LEARNED) {
        learned.setup(ctx, ctx->mpnn, A, B, alpha, beta);
        learned_on = true;
    }
    const bool no_fallback = opts && opts->no_fallback;
    // host -> device (column-major host, row-major device): b and the start x
    std::vector<double> hb((size_t)n * m), hx((size_t)n * m);
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < m; ++c) {
            hb[i * m + c] = b[(int64_t)c * n + i];
            hx[i * m + c] = x[(int64_t)c * n + i];
        }
    const size_t blk = (size_t)n * m * 8;
    DevBuf dB(blk), dX(blk), dR(blk), dZ(blk), dP(blk), dQ(blk);
    h2d(ctx, dZ.p, hb.data(), blk);
    h2d(ctx, dP.p, hx.data(), blk);
    // b and x0 with the Dirichlet rows cleared (those rows are not solved
    // for), in the solve's row order
    if (ro) {
        k_mask_rows<<<nblocks(n * m, 256), 256, 0, ctx->stream>>>(n, m, dmp, dZ.as<double>(), dR.as<double>());
        k_mask_rows<<<nblocks(n * m, 256), 256, 0, ctx->stream>>>(n, m, dmp, dP.as<double>(), dQ.as<double>());
        k_perm_rows<<<nblocks(n * m, 256), 256, 0, ctx->stream>>>(n, m, ro->perm.as<int32_t>(), dR.as<double>(),
                                                                 dB.as<double>());
        k_perm_rows<<<nblocks(n * m, 256), 256, 0, ctx->stream>>>(n, m, ro->perm.as<int32_t>(), dQ.as<double>(),
                                                                 dX.as<double>());
    } else {
        k_mask_rows<<<nblocks(n * m, 256), 256, 0, ctx->stream>>>(n, m, dmp, dZ.as<double>(), dB.as<double>());
        k_mask_rows<<<nblocks(n * m, 256), 256, 0, ctx->stream>>>(n, m, dmp, dP.as<double>(), dX.as<double>());
    }
    LH_CHECK_LAUNCH();
    if (singular) {  // b onto the complement of the constants
        L.coldot(dB.as<double>(), nullptr, ptmp, sums);
        k_sub_colmean<<<nblocks(n * m, 256), 256, 0, ctx->stream>>>(n, m, dB.as<double>(), sums);
        LH_CHECK_LAUNCH();
    }
    // r = b - K x0 (the caller's x warm-starts the free rows)
    L.resid(dX.as<double>(), dB.as<double>(), dR.as<double>());
    if (singular) {
        L.coldot(dR.as<double>(), nullptr, ptmp, sums);
        k_sub_colmean<<<nblocks(n * m, 256), 256, 0, ctx->stream>>>(n, m, dR.as<double>(), sums);
        LH_CHECK_LAUNCH();
    }
    std::vector<double> hbb(m), htol(m);
    L.coldot(dB.as<double>(), dB.as<double>(), ptmp, sums);
    d2h(ctx, hbb.data(), sums, (size_t)m * 8);
    // a column with b = 0 is never active (its solution is x = 0)
    for (int c = 0; c < m; ++c) htol[c] = hbb[c] > 0.0 ? rtol * rtol * hbb[c] : HUGE_VAL;
    h2d(ctx, tol2, htol.data(), (size_t)m * 8);
    // partials of r_0.r_0 into parity 0
    L.coldot_part(dR.as<double>(), dR.as<double>(), s.rr[0]);
    // z = M^-1 r with the learned network (Dirichlet rows kept at zero)
    auto apply_learned = [&]() {
        HIPX(hipMemcpyAsync(dZ.p, dR.p, blk, hipMemcpyDeviceToDevice, ctx->stream));
        learned.apply(dZ.as<double>(), m, m);
        if (dmp) k_mask_rows<<<nblocks(n * m, 256), 256, 0, ctx->stream>>>(n, m, dmp, dZ.as<double>(), dZ.as<double>());
        LH_CHECK_LAUNCH();
    };
    // a first direction at iteration k: z = M^-1 r and the partials of r.z
    // into parity k % 2, P = Q = 0 (so p = z, q = K z whatever beta is)
    auto start_direction = [&](bool jac, int64_t k) {
        if (jac && pdeg > 1) {
            L.cheb(pdeg, cheb_amin, cheb_bmax, dinv.as<double>(), dR.as<double>(), dZ.as<double>(), dT.as<double>(),
                   nullptr);
        } else if (jac) {
            HIPX(hipMemcpyAsync(dZ.p, dR.p, blk, hipMemcpyDeviceToDevice, ctx->stream));
            rowscale<double>(ctx, n, m, dZ.as<double>(), m, dinv.as<double>());
        } else {
            apply_learned();
        }
        L.rz(dR.as<double>(), dZ.as<double>(), s.rz[k & 1], nullptr);
        HIPX(hipMemsetAsync(dP.p, 0, blk, ctx->stream));
        HIPX(hipMemsetAsync(dQ.p, 0, blk, ctx->stream));
    };
    start_direction(!learned_on, 0);
    int64_t enq = 0;
    int done = 0, iters = 0;
    double worst = 0.0, prev_worst = 1.0;
    int slow = 0;
    int* hstat = static_cast<int*>(ctx->host_small);
    double* hrr_pin = reinterpret_cast<double*>(static_cast<char*>(ctx->host_small) + 64);
    while (!done && enq < max_iter) {
        const int chunk = (int)std::min<int64_t>(learned_on ? 4 : 16, max_iter - enq);
        for (int kk = 0; kk < chunk; ++kk) {
            const int par = (int)((enq + kk) & 1);
            L.dir(par, singular, dZ.as<double>(), dP.as<double>(), dQ.as<double>(), dR.as<double>(), s);
            L.update(par, singular, !learned_on && pdeg == 1, dinv.as<double>(), dP.as<double>(), dQ.as<double>(),
                     dX.as<double>(), dR.as<double>(), dZ.as<double>(), s);
            if (learned_on) {
                apply_learned();
                L.rz(dR.as<double>(), dZ.as<double>(), s.rz[par ^ 1], s.status);
            } else if (pdeg > 1) {
                L.cheb(pdeg, cheb_amin, cheb_bmax, dinv.as<double>(), dR.as<double>(), dZ.as<double>(),
                       dT.as<double>(), s.status);
                L.rz(dR.as<double>(), dZ.as<double>(), s.rz[par ^ 1], s.status);
            }
            st.spmm_calls++;
        }
        enq += chunk;
        HIPX(hipMemcpyAsync(hstat, s.status, 12, hipMemcpyDeviceToHost, ctx->stream));
        HIPX(hipMemcpyAsync(hrr_pin, rrout, (size_t)m * 8, hipMemcpyDeviceToHost, ctx->stream));
        HIPX(hipStreamSynchronize(ctx->stream));
        done = hstat[0];
        iters = hstat[1];
        worst = 0.0;
        for (int c = 0; c < m; ++c)
            if (hbb[c] > 0.0) worst = std::max(worst, std::sqrt(hrr_pin[c] / hbb[c]));
        // learned preconditioner: back to Jacobi (and a CG restart) when the
        // residual falls by less than half per iteration over three checks
        // running
        if (learned_on && !no_fallback && !done) {
            const double per_it = prev_worst > 0.0 ? std::pow(worst / prev_worst, 1.0 / chunk) : 0.0;
            slow = (per_it > 0.5) ? slow + 1 : 0;
            if (slow >= 3) {
                learned_on = false;
                st.fallback_used = 1;
                start_direction(true, enq);
            }
        }
        prev_worst = worst;
        if (verbose && (enq % 256 == 0)) fprintf(stderr, "[pcg] it %d worst %.3e\n", iters, worst);
    }
    if (singular) {
        L.coldot(dX.as<double>(), nullptr, ptmp, sums);
        k_sub_colmean<<<nblocks(n * m, 256), 256, 0, ctx->stream>>>(n, m, dX.as<double>(), sums);
        LH_CHECK_LAUNCH();
    }
    if (ro) {
        ro->unpermute(ctx, dX.as<double>(), m, dZ.as<double>(), m, n, m);
        d2h(ctx, hb.data(), dZ.p, blk);
    } else {
        d2h(ctx, hb.data(), dX.p, blk);
    }
    int nconv = 0;
    for (int c = 0; c < m; ++c) nconv += (hbb[c] == 0.0 || hrr_pin[c] <= htol[c] || done) ? 1 : 0;
    for (int64_t i = 0; i < n; ++i) {
        const bool dir = any_dirichlet && dmask[i];
        for (int c = 0; c < m; ++c)
            if (!dir) x[(int64_t)c * n + i] = hbb[c] > 0.0 ? hb[i * m + c] : 0.0;  // b = 0: x = 0
    }
    st.iterations = iters;
    st.converged = nconv;
    st.max_resid = worst;
    st.t_total_us = now_us() - t0;
    if (stats) *stats = st;
    if (nconv < m) {
        ctx->last_error = "PCG did not converge within max_iter";
        return LH_ERR_NOCONV;
    }
    LH_API_END(ctx)
}

}  // extern "C"
