// Chebyshev-accelerated Jacobi preconditioner for K = A - sigma B.
//
// z = p(D^-1 K) D^-1 r is the result of `degree` steps of the Chebyshev
// semi-iterative method for K z = r (Jacobi-preconditioned, z_0 = 0) on the
// interval [lmax / ratio, lmax] of D^-1 K; lmax comes from power iteration.
// On that interval x p(x) stays within 1 +- 1/T_d((ratio+1)/(ratio-1)); below
// it x p(x) rises monotonically from 0, with slope about 2 d^2 / lmax, which
// is what cuts the block eigensolver's condition number ~ 2 d^2 times
// compared with plain Jacobi (degree 1).  The operator is SPD for SPD K.
//
// Each step is one fused kernel: gather-SpMM of the current direction with
// the values of K, then the residual, direction and iterate updates of the
// same row (the direction is double-buffered, so rows never read a neighbour
// that was already overwritten).
#include "lh_dense.hpp"

#include <cmath>

namespace lh {

namespace {

template <typename VT>
__global__ void k_shift_vals(int64_t nnz, const VT* __restrict__ av, const VT* __restrict__ bv, double alpha,
                             double beta, double* __restrict__ kv) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nnz) return;
    kv[s] = beta * (double)av[s] + (bv ? alpha * (double)bv[s] : 0.0);
}
template <typename VT>
__global__ void k_shift_diag(int64_t n, const int32_t* __restrict__ diagpos, const VT* __restrict__ d, double alpha,
                             double* __restrict__ kv, int* __restrict__ missing) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (d[i] == VT(0)) return;
    if (diagpos[i] < 0) { atomicOr(missing, 1); return; }
    kv[diagpos[i]] += alpha * (double)d[i];
}
__global__ void k_dinv_from(int64_t n, const int32_t* __restrict__ diagpos, const double* __restrict__ kv,
                            double* __restrict__ dinv) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d = diagpos[i] >= 0 ? kv[diagpos[i]] : 0.0;
    dinv[i] = d != 0.0 ? 1.0 / d : 0.0;
}

// b = W (FP32 copy), x_1 = alpha D^-1 b, x_0 = 0
template <typename WT>
__global__ void k_cheb_init(int64_t n, int m, const WT* __restrict__ W, int ldw, const float* __restrict__ dinv,
                            float alpha, float* __restrict__ bvec, float* __restrict__ x1, float* __restrict__ x0) {
    int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n * m) return;
    int64_t i = gid / m;
    int c = (int)(gid - i * m);
    const float r = (float)W[i * ldw + c];
    bvec[gid] = r;
    x1[gid] = alpha * dinv[i] * r;
    x0[gid] = 0.0f;
}

// x_{k+1} = omega (x_k - x_{k-1} + alpha D^-1 (b - K x_k)) + x_{k-1}, written
// over x_{k-1} (row-local), x_k gathered
template <int LPR>
__global__ void __launch_bounds__(256)
k_cheb_step(int64_t n, int m, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cols,
            const float* __restrict__ kv, const float* __restrict__ dinv, const float* __restrict__ bvec,
            const float* __restrict__ xk, float* __restrict__ xprev, float alpha, float omega) {
    const int64_t gid = xcd_block() * blockDim.x + threadIdx.x;
    const int64_t row = gid / LPR;
    const int lane = (int)(gid % LPR);
    if (row >= n) return;
    const int32_t r0 = rowptr[row], r1 = rowptr[row + 1];
    const float ad = alpha * dinv[row];
    for (int c = lane; c < m; c += LPR) {
        float acc = 0.0f;
        for (int32_t s = r0; s < r1; ++s) acc += kv[s] * xk[(int64_t)cols[s] * m + c];
        const int64_t o = row * m + c;
        const float xp = xprev[o];
        xprev[o] = omega * (xk[o] - xp + ad * (bvec[o] - acc)) + xp;
    }
}

// vectorised form: 4 columns per lane, 16 lanes per row for m = 64
template <int LPR>
__global__ void __launch_bounds__(256)
k_cheb_step_v(int64_t n, int m, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cols,
              const float* __restrict__ kv, const float* __restrict__ dinv, const float* __restrict__ bvec,
              const float* __restrict__ xk, float* __restrict__ xprev, float alpha, float omega) {
    const int64_t gid = xcd_block() * blockDim.x + threadIdx.x;
    const int64_t row = gid / LPR;
    const int c0 = (int)(gid % LPR) * 4;
    if (row >= n || c0 >= m) return;
    const float4 acc = row_gather<float, float>(cols, kv, rowptr[row], rowptr[row + 1], xk, m, c0);
    const float ad = alpha * dinv[row];
    const int64_t o = row * m + c0;
    const float4 xkv = *reinterpret_cast<const float4*>(xk + o);
    const float4 bv = *reinterpret_cast<const float4*>(bvec + o);
    float4 xp = *reinterpret_cast<const float4*>(xprev + o);
    xp.x = omega * (xkv.x - xp.x + ad * (bv.x - acc.x)) + xp.x;
    xp.y = omega * (xkv.y - xp.y + ad * (bv.y - acc.y)) + xp.y;
    xp.z = omega * (xkv.z - xp.z + ad * (bv.z - acc.z)) + xp.z;
    xp.w = omega * (xkv.w - xp.w + ad * (bv.w - acc.w)) + xp.w;
    *reinterpret_cast<float4*>(xprev + o) = xp;
}

// ---- FP16 storage, FP32 arithmetic --------------------------------------
// The block is scaled per column to RMS 2^-4 before the first step, so that
// the smooth components, which the polynomial amplifies at most ~2 d^2 /
// lmax times (times D^-1), stay far below the FP16 limit; stores saturate
// at +-65504 in any case.  The scales are undone on output.
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float float8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ float8 h2f(half8 h) { return __builtin_convertvector(h, float8); }
__device__ __forceinline__ _Float16 f2h_sat(float v) {
    return (_Float16)fminf(fmaxf(v, -65504.0f), 65504.0f);
}
__device__ __forceinline__ half8 f2h_sat(float8 v) {
    half8 h;
#pragma unroll
    for (int q = 0; q < 8; ++q) h[q] = f2h_sat(v[q]);
    return h;
}

// scale[c] = target / ||w_c|| (1 for a zero column), inv[c] = 1 / scale[c]
__global__ void k_col_scales(int m, const double* __restrict__ colnorm2, double target, float* __restrict__ scale,
                             double* __restrict__ inv) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m) return;
    const double q = colnorm2[c];
    const double sc = q > 0.0 ? target / sqrt(q) : 1.0;
    scale[c] = (float)sc;
    inv[c] = 1.0 / (double)(float)sc;
}

// b = s_c W, x_1 = alpha D^-1 b, x_0 = 0 (FP16)
template <typename WT>
__global__ void k_cheb_init_h(int64_t n, int m, const WT* __restrict__ W, int ldw, const float* __restrict__ dinv,
                              const float* __restrict__ scale, float alpha, _Float16* __restrict__ bvec,
                              _Float16* __restrict__ x1, _Float16* __restrict__ x0) {
    int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n * m) return;
    int64_t i = gid / m;
    int c = (int)(gid - i * m);
    const float r = (float)W[i * ldw + c] * scale[c];
    bvec[gid] = f2h_sat(r);
    x1[gid] = f2h_sat(alpha * dinv[i] * r);
    x0[gid] = (_Float16)0.0f;
}

// one step, 8 columns (16 bytes) per lane; requires m % 8 == 0.  The row's
// own x_k, b and x_{k-1} are loaded first, so that they are in flight with the
// gathers; the gathers go four at a time (a partial last group repeats the
// row's own x_k with weight 0)
template <int LPR>
__global__ void __launch_bounds__(256)
k_cheb_step_h(int64_t n, int m, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cols,
              const float* __restrict__ kv, const float* __restrict__ dinv, const _Float16* __restrict__ bvec,
              const _Float16* __restrict__ xk, _Float16* __restrict__ xprev, float alpha, float omega) {
    const int64_t gid = xcd_block() * blockDim.x + threadIdx.x;
    const int64_t row = gid / LPR;
    const int c0 = (int)(gid % LPR) * 8;
    if (row >= n || c0 >= m) return;
    const int64_t o = row * m + c0;
    const half8 hk = *reinterpret_cast<const half8*>(xk + o);
    const half8 hb = *reinterpret_cast<const half8*>(bvec + o);
    const half8 hp = *reinterpret_cast<const half8*>(xprev + o);
    const float ad = alpha * dinv[row];
    const int32_t r0 = rowptr[row], r1 = rowptr[row + 1];
    float8 a0 = 0.0f, a1 = 0.0f;
    for (int32_t s = r0; s < r1; s += 4) {
        int32_t j[4];
        float w[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const bool in = s + q < r1;
            j[q] = in ? cols[s + q] : (int32_t)row;
            w[q] = in ? kv[s + q] : 0.0f;
        }
        half8 x[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) x[q] = *reinterpret_cast<const half8*>(xk + (int64_t)j[q] * m + c0);
        a0 += w[0] * h2f(x[0]);
        a1 += w[1] * h2f(x[1]);
        a0 += w[2] * h2f(x[2]);
        a1 += w[3] * h2f(x[3]);
    }
    a0 += a1;
    const float8 xkv = h2f(hk), bv = h2f(hb), xp = h2f(hp);
    const float8 r = omega * (xkv - xp + ad * (bv - a0)) + xp;
    *reinterpret_cast<half8*>(xprev + o) = f2h_sat(r);
}

// ELL form of the step: every row has W slots (W % 8 == 0, row-major; a
// padding slot points at the row itself with weight 0), so the slots' columns
// and weights do not wait for the row pointers, and eight gathers go out at
// once.  W == 8 (surface meshes): the workgroup's rows' slots are staged in
// LDS by one coalesced load instead of being fetched by every lane of a row.
template <int LPR, bool W8>
__global__ void __launch_bounds__(256)
k_cheb_step_e(int64_t n, int m, int W, const int32_t* __restrict__ ecols, const float* __restrict__ ekv,
              const float* __restrict__ dinv, const _Float16* __restrict__ bvec, const _Float16* __restrict__ xk,
              _Float16* __restrict__ xprev, float alpha, float omega) {
    constexpr int kRows = 256 / LPR;  // rows of a workgroup
    __shared__ __align__(16) int32_t s_cols[W8 ? kRows * 8 : 4];
    __shared__ __align__(16) float s_kv[W8 ? kRows * 8 : 4];
    const int64_t blk = xcd_block();
    const int64_t gid = blk * blockDim.x + threadIdx.x;
    const int64_t row = gid / LPR;
    const int c0 = (int)(gid % LPR) * 8;
    if (W8) {
        // rows blk * kRows ... (+kRows): kRows * 8 slots, contiguous
        const int64_t base = blk * kRows * 8;
        for (int t = threadIdx.x; t < kRows * 8; t += 256) {
            const bool in = base + t < n * 8;
            s_cols[t] = in ? ecols[base + t] : 0;
            s_kv[t] = in ? ekv[base + t] : 0.0f;
        }
        __syncthreads();
    }
    if (row >= n || c0 >= m) return;
    const int64_t o = row * m + c0;
    const half8 hk = *reinterpret_cast<const half8*>(xk + o);
    // b and x_{k-1} are streamed once per step: non-temporal
    const half8 hb = __builtin_nontemporal_load(reinterpret_cast<const half8*>(bvec + o));
    const half8 hp = __builtin_nontemporal_load(reinterpret_cast<const half8*>(xprev + o));
    const float ad = alpha * dinv[row];
    const int32_t* cr = W8 ? s_cols + (threadIdx.x / LPR) * 8 : ecols + row * W;
    const float* wr = W8 ? s_kv + (threadIdx.x / LPR) * 8 : ekv + row * W;
    float8 a0 = 0.0f, a1 = 0.0f;
    for (int s = 0; s < (W8 ? 8 : W); s += 8) {
        const int4 ja = *reinterpret_cast<const int4*>(cr + s);
        const int4 jb = *reinterpret_cast<const int4*>(cr + s + 4);
        const float4 wa = *reinterpret_cast<const float4*>(wr + s);
        const float4 wb = *reinterpret_cast<const float4*>(wr + s + 4);
        const int32_t j[8] = {ja.x, ja.y, ja.z, ja.w, jb.x, jb.y, jb.z, jb.w};
        const float w[8] = {wa.x, wa.y, wa.z, wa.w, wb.x, wb.y, wb.z, wb.w};
        half8 x[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) x[q] = *reinterpret_cast<const half8*>(xk + (int64_t)j[q] * m + c0);
#pragma unroll
        for (int q = 0; q < 8; q += 2) {
            a0 += w[q] * h2f(x[q]);
            a1 += w[q + 1] * h2f(x[q + 1]);
        }
    }
    a0 += a1;
    const float8 xkv = h2f(hk), bv = h2f(hb), xp = h2f(hp);
    const float8 r = omega * (xkv - xp + ad * (bv - a0)) + xp;
    __builtin_nontemporal_store(f2h_sat(r), reinterpret_cast<half8*>(xprev + o));
}

// longest row of a pattern (into *w, which starts at 0)
__global__ void k_max_rowlen(int64_t n, const int32_t* __restrict__ rowptr, int* __restrict__ w) {
    __shared__ int red[256];
    int best = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        best = max(best, rowptr[i + 1] - rowptr[i]);
    red[threadIdx.x] = best;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(w, red[0]);
}

// CSR -> ELL with W slots per row (padding: the row itself, weight 0)
__global__ void k_to_ell(int64_t n, int W, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cols,
                         const float* __restrict__ kv, int32_t* __restrict__ ecols, float* __restrict__ ekv) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t r0 = rowptr[i], len = rowptr[i + 1] - r0;
    for (int q = 0; q < W; ++q) {
        const bool in = q < len;
        ecols[i * W + q] = in ? cols[r0 + q] : (int32_t)i;
        ekv[i * W + q] = in ? kv[r0 + q] : 0.0f;
    }
}

template <typename WT>
__global__ void k_cheb_out_h(int64_t n, int m, const _Float16* __restrict__ x, const double* __restrict__ inv,
                             WT* __restrict__ W, int ldw) {
    int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n * m) return;
    int64_t i = gid / m;
    int c = (int)(gid - i * m);
    W[i * ldw + c] = (WT)((double)(float)x[gid] * inv[c]);
}

template <typename WT>
__global__ void k_cheb_out(int64_t n, int m, const float* __restrict__ x, WT* __restrict__ W, int ldw) {
    int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n * m) return;
    int64_t i = gid / m;
    int c = (int)(gid - i * m);
    W[i * ldw + c] = (WT)x[gid];
}

__global__ void k_to_f32(int64_t cnt, const double* __restrict__ s, float* __restrict__ d) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) d[i] = (float)s[i];
}

__global__ void k_rayleigh_parts(int64_t n, const double* __restrict__ x, const double* __restrict__ kx,
                                 const int32_t* __restrict__ diagpos, const double* __restrict__ kv,
                                 double* __restrict__ part) {
    // part[block] = (x^T K x, x^T D x, x^T x)
    __shared__ double s[3][256];
    double a = 0.0, b = 0.0, c = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double d = diagpos[i] >= 0 ? kv[diagpos[i]] : 0.0;
        a += x[i] * kx[i];
        b += x[i] * d * x[i];
        c += x[i] * x[i];
    }
    s[0][threadIdx.x] = a;
    s[1][threadIdx.x] = b;
    s[2][threadIdx.x] = c;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < (unsigned)off)
            for (int q = 0; q < 3; ++q) s[q][threadIdx.x] += s[q][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; ++q) part[blockIdx.x * 3 + q] = s[q][0];
}

// x <- dinv * kx  (power step for D^-1 K); x <- s x
__global__ void k_power_next(int64_t n, const double* __restrict__ kx, const double* __restrict__ dinv,
                             double* __restrict__ x) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = dinv[i] * kx[i];
}
__global__ void k_scale_vec(int64_t n, double* __restrict__ x, double s) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] *= s;
}

}  // namespace

void form_shifted(lh_ctx* ctx, const lh_csr* A, const lh_csr* B, double alpha, double beta, DevBuf& kv) {
    const int64_t n = A->n(), nnz = A->nnz();
    kv.ensure(nnz * 8);
    const bool use_b = (B != nullptr) && alpha != 0.0;
    if (use_b && !B->is_diag && B->pat != A->pat)
        LH_THROW(LH_ERR_ARG, "B must be diagonal or share A's sparsity pattern");
    DevBuf missing(8);
    memset_d(ctx, missing.p, 0, 4);
    if (A->dtype == LH_F64) {
        k_shift_vals<double><<<nblocks(nnz, 256), 256, 0, ctx->stream>>>(
            nnz, A->vals.as<double>(), (use_b && !B->is_diag) ? B->vals.as<double>() : nullptr, alpha, beta,
            kv.as<double>());
        if (use_b && B->is_diag)
            k_shift_diag<double><<<nblocks(n, 256), 256, 0, ctx->stream>>>(
                n, A->pat->diagpos.as<int32_t>(), B->dense_diag.as<double>(), alpha, kv.as<double>(), missing.as<int>());
    } else {
        k_shift_vals<float><<<nblocks(nnz, 256), 256, 0, ctx->stream>>>(
            nnz, A->vals.as<float>(), (use_b && !B->is_diag) ? B->vals.as<float>() : nullptr, alpha, beta,
            kv.as<double>());
        if (use_b && B->is_diag)
            k_shift_diag<float><<<nblocks(n, 256), 256, 0, ctx->stream>>>(
                n, A->pat->diagpos.as<int32_t>(), B->dense_diag.as<float>(), alpha, kv.as<double>(), missing.as<int>());
    }
    LH_CHECK_LAUNCH();
    int miss = 0;
    d2h(ctx, &miss, missing.p, 4);
    if (miss) LH_THROW(LH_ERR_ARG, "B has diagonal entries outside A's pattern");
}

void ChebPrecond::setup(lh_ctx* c, const lh_csr* A, const lh_csr* B, double sigma, int deg, double ratio,
                        int64_t n_, int mmax_, bool half_storage) {
    ctx = c;
    mmax = mmax_;
    half = half_storage;
    if (half && mmax > 512) LH_THROW(LH_ERR_ARG, "FP16 Chebyshev storage takes at most 512 block columns (%d)", mmax);
    pat = A->pat;
    n = n_;
    degree = deg;
    form_shifted(ctx, A, B, -sigma, 1.0, kv);
    dinv.alloc(n * 8);
    k_dinv_from<<<nblocks(n, 256), 256, 0, ctx->stream>>>(n, pat->diagpos.as<int32_t>(), kv.as<double>(),
                                                          dinv.as<double>());
    LH_CHECK_LAUNCH();
    // power iteration for lmax of D^-1 K (Rayleigh quotient x^T K x / x^T D x)
    DevBuf x(n * 8), kx(n * 8), part(256 * 3 * 8);
    randn<double>(ctx, n, 1, x.as<double>(), 1, 0x5eed);
    double lmax = 2.0, hp[256 * 3];
    lh_csr Kv;
    Kv.ctx = ctx;
    Kv.pat = pat;
    Kv.dtype = LH_F64;
    Kv.vals.p = kv.p;
    for (int it = 0; it < 40; ++it) {
        spmm_ld<double, double>(ctx, &Kv, x.as<double>(), kx.as<double>(), 1, 1, 1);
        k_rayleigh_parts<<<256, 256, 0, ctx->stream>>>(n, x.as<double>(), kx.as<double>(), pat->diagpos.as<int32_t>(),
                                                      kv.as<double>(), part.as<double>());
        LH_CHECK_LAUNCH();
        d2h(ctx, hp, part.p, sizeof(hp));
        double a = 0, b = 0, cc = 0;
        for (int q = 0; q < 256; ++q) { a += hp[3 * q]; b += hp[3 * q + 1]; cc += hp[3 * q + 2]; }
        lmax = (b > 0) ? a / b : 2.0;
        // next x = D^-1 K x, normalised
        k_power_next<<<nblocks(n, 256), 256, 0, ctx->stream>>>(n, kx.as<double>(), dinv.as<double>(), x.as<double>());
        LH_CHECK_LAUNCH();
        k_rayleigh_parts<<<256, 256, 0, ctx->stream>>>(n, x.as<double>(), x.as<double>(), pat->diagpos.as<int32_t>(),
                                                      kv.as<double>(), part.as<double>());
        d2h(ctx, hp, part.p, sizeof(hp));
        double nn = 0;
        for (int q = 0; q < 256; ++q) nn += hp[3 * q + 2];
        const double sc = nn > 0 ? 1.0 / std::sqrt(nn) : 1.0;
        k_scale_vec<<<nblocks(n, 256), 256, 0, ctx->stream>>>(n, x.as<double>(), sc);
        LH_CHECK_LAUNCH();
        (void)cc;
    }
    Kv.vals.p = nullptr;
    bmax = 1.1 * lmax;  // power iteration approaches lmax from below
    amin = bmax / (ratio > 1.0 ? ratio : 64.0);
    // FP32 copies of K and D^-1 and the FP32 work blocks: the
    // preconditioner only shapes the search directions, so its rounding does
    // not limit the accuracy of the Ritz pairs
    const int64_t nnz = pat->nnz;
    kv32.alloc(nnz * 4);
    dinv32.alloc(n * 4);
    k_to_f32<<<nblocks(nnz, 256), 256, 0, ctx->stream>>>(nnz, kv.as<double>(), kv32.as<float>());
    k_to_f32<<<nblocks(n, 256), 256, 0, ctx->stream>>>(n, dinv.as<double>(), dinv32.as<float>());
    LH_CHECK_LAUNCH();
    // ELL copy for the FP16 step when the rows are short (surface and
    // volume meshes: 7 and ~15 slots)
    ell_w = 0;
    if (half) {
        DevBuf wmax(8);
        memset_d(ctx, wmax.p, 0, 4);
        k_max_rowlen<<<256, 256, 0, ctx->stream>>>(n, pat->rowptr.as<int32_t>(), wmax.as<int>());
        LH_CHECK_LAUNCH();
        int w = 0;
        d2h(ctx, &w, wmax.p, 4);
        w = (w + 7) / 8 * 8;
        if (w > 0 && w <= kEllMaxW) {
            ell_w = w;
            ecols.alloc((size_t)n * w * 4);
            ekv.alloc((size_t)n * w * 4);
            k_to_ell<<<nblocks(n, 256), 256, 0, ctx->stream>>>(n, w, pat->rowptr.as<int32_t>(), pat->cols.as<int32_t>(),
                                                              kv32.as<float>(), ecols.as<int32_t>(), ekv.as<float>());
            LH_CHECK_LAUNCH();
        }
    }
    const size_t es = half ? 2 : 4;
    R.alloc(n * mmax * es);
    D0.alloc(n * mmax * es);
    D1.alloc(n * mmax * es);
}

// one Chebyshev step x_prev <- step(x_k, x_prev) on an n x m block
void ChebPrecond::step(int m, const void* bvec, const void* xk, void* xp, double omega) {
    const int32_t* rp = pat->rowptr.as<int32_t>();
    const int32_t* cl = pat->cols.as<int32_t>();
    const float* kk = kv32.as<float>();
    const float* dd = dinv32.as<float>();
    const float fa = (float)(2.0 / (bmax + amin)), fo = (float)omega;
    if (half) {
        auto* b = static_cast<const _Float16*>(bvec);
        auto* x = static_cast<const _Float16*>(xk);
        auto* y = static_cast<_Float16*>(xp);
        const int vl = vec_lanes(m, 8);
        const unsigned nb = nblocks_xcd(n * vl, 256);
        if (ell_w > 0) {
            const int32_t* ec = ecols.as<int32_t>();
            const float* ek = ekv.as<float>();
            const int W = ell_w;
#define LH_CHEB_E(L)                                                                                     \
    do {                                                                                                 \
        if (W == 8) k_cheb_step_e<L, true><<<nb, 256, 0, ctx->stream>>>(n, m, W, ec, ek, dd, b, x, y, fa, fo); \
        else k_cheb_step_e<L, false><<<nb, 256, 0, ctx->stream>>>(n, m, W, ec, ek, dd, b, x, y, fa, fo);       \
    } while (0)
            switch (vl) {
                case 1: LH_CHEB_E(1); break;
                case 2: LH_CHEB_E(2); break;
                case 4: LH_CHEB_E(4); break;
                case 8: LH_CHEB_E(8); break;
                case 16: LH_CHEB_E(16); break;
                case 32: LH_CHEB_E(32); break;
                default: LH_CHEB_E(64); break;  // vl <= 64: checked at setup
            }
#undef LH_CHEB_E
            LH_CHECK_LAUNCH();
            return;
        }
        switch (vl) {
            case 1: k_cheb_step_h<1><<<nb, 256, 0, ctx->stream>>>(n, m, rp, cl, kk, dd, b, x, y, fa, fo); break;
            case 2: k_cheb_step_h<2><<<nb, 256, 0, ctx->stream>>>(n, m, rp, cl, kk, dd, b, x, y, fa, fo); break;
            case 4: k_cheb_step_h<4><<<nb, 256, 0, ctx->stream>>>(n, m, rp, cl, kk, dd, b, x, y, fa, fo); break;
            case 8: k_cheb_step_h<8><<<nb, 256, 0, ctx->stream>>>(n, m, rp, cl, kk, dd, b, x, y, fa, fo); break;
            case 16: k_cheb_step_h<16><<<nb, 256, 0, ctx->stream>>>(n, m, rp, cl, kk, dd, b, x, y, fa, fo); break;
            case 32: k_cheb_step_h<32><<<nb, 256, 0, ctx->stream>>>(n, m, rp, cl, kk, dd, b, x, y, fa, fo); break;
            default: k_cheb_step_h<64><<<nb, 256, 0, ctx->stream>>>(n, m, rp, cl, kk, dd, b, x, y, fa, fo); break;
        }
        LH_CHECK_LAUNCH();
        return;
    }
    auto* b = static_cast<const float*>(bvec);
    auto* x = static_cast<const float*>(xk);
    auto* y = static_cast<float*>(xp);
    const bool vec = (m % 4 == 0) && vec_lanes(m, 4) >= 4 && vec_lanes(m, 4) <= 64;
    if (vec) {
        const int vl = vec_lanes(m, 4);
        const unsigned nbv = nblocks_xcd(n * vl, 256);
        switch (vl) {
            case 4: k_cheb_step_v<4><<<nbv, 256, 0, ctx->stream>>>(n, m, rp, cl, kk, dd, b, x, y, fa, fo); break;
            case 8: k_cheb_step_v<8><<<nbv, 256, 0, ctx->stream>>>(n, m, rp, cl, kk, dd, b, x, y, fa, fo); break;
            case 16: k_cheb_step_v<16><<<nbv, 256, 0, ctx->stream>>>(n, m, rp, cl, kk, dd, b, x, y, fa, fo); break;
            case 32: k_cheb_step_v<32><<<nbv, 256, 0, ctx->stream>>>(n, m, rp, cl, kk, dd, b, x, y, fa, fo); break;
            default: k_cheb_step_v<64><<<nbv, 256, 0, ctx->stream>>>(n, m, rp, cl, kk, dd, b, x, y, fa, fo); break;
        }
    } else {
        const int lpr = m > 32 ? 64 : (m > 16 ? 32 : 16);
        const unsigned nb = nblocks_xcd(n * lpr, 256);
        if (lpr == 64) k_cheb_step<64><<<nb, 256, 0, ctx->stream>>>(n, m, rp, cl, kk, dd, b, x, y, fa, fo);
        else if (lpr == 32) k_cheb_step<32><<<nb, 256, 0, ctx->stream>>>(n, m, rp, cl, kk, dd, b, x, y, fa, fo);
        else k_cheb_step<16><<<nb, 256, 0, ctx->stream>>>(n, m, rp, cl, kk, dd, b, x, y, fa, fo);
    }
    LH_CHECK_LAUNCH();
}

// b, x_1, x_0 from W; returns the (x_k, x_prev) buffers
template <typename WT>
void ChebPrecond::init(const WT* W, int ldw, int m, const double* colnorm2, void** xk, void** xp) {
    const float fa = (float)(2.0 / (bmax + amin));
    if (half) {
        if (!colnorm2) LH_THROW(LH_ERR_ARG, "FP16 Chebyshev storage needs the column norms of the block");
        sc.ensure(m * 16);
        float* scale = sc.as<float>();
        double* inv = reinterpret_cast<double*>(sc.as<char>() + 8 * m);
        k_col_scales<<<nblocks(m, 256), 256, 0, ctx->stream>>>(m, colnorm2, std::sqrt((double)n) / 16.0, scale, inv);
        k_cheb_init_h<<<nblocks(n * m, 256), 256, 0, ctx->stream>>>(n, m, W, ldw, dinv32.as<float>(), scale, fa,
                                                                    R.as<_Float16>(), D0.as<_Float16>(),
                                                                    D1.as<_Float16>());
    } else {
        k_cheb_init<<<nblocks(n * m, 256), 256, 0, ctx->stream>>>(n, m, W, ldw, dinv32.as<float>(), fa, R.as<float>(),
                                                                  D0.as<float>(), D1.as<float>());
    }
    LH_CHECK_LAUNCH();
    *xk = D0.p;
    *xp = D1.p;
}

template <typename WT>
void ChebPrecond::apply(WT* W, int ldw, int m, const double* colnorm2) {
    // Chebyshev semi-iteration (Golub-Varga two-step form) for K x = w,
    // Jacobi-scaled, on [amin, bmax]
    if (m > mmax) LH_THROW(LH_ERR_ARG, "Chebyshev block width %d exceeds %d", m, mmax);
    if (half && m % 8 != 0) LH_THROW(LH_ERR_ARG, "FP16 Chebyshev storage needs a block width divisible by 8");
    const double mu = (bmax - amin) / (bmax + amin);
    void *xk, *xp;
    init(W, ldw, m, colnorm2, &xk, &xp);
    double omega = 1.0;
    for (int k = 1; k < degree; ++k) {
        omega = (k == 1) ? 1.0 / (1.0 - 0.5 * mu * mu) : 1.0 / (1.0 - 0.25 * mu * mu * omega);
        step(m, R.p, xk, xp, omega);
        std::swap(xk, xp);
    }
    if (half) {
        const double* inv = reinterpret_cast<const double*>(sc.as<char>() + 8 * m);
        k_cheb_out_h<<<nblocks(n * m, 256), 256, 0, ctx->stream>>>(n, m, static_cast<const _Float16*>(xk), inv, W,
                                                                   ldw);
    } else {
        k_cheb_out<<<nblocks(n * m, 256), 256, 0, ctx->stream>>>(n, m, static_cast<const float*>(xk), W, ldw);
    }
    LH_CHECK_LAUNCH();
}
template void ChebPrecond::apply<double>(double*, int, int, const double*);
template void ChebPrecond::apply<float>(float*, int, int, const double*);

double ChebPrecond::time_step(int m, int iters) {
    // random block, then `iters` back-to-back steps between HIP events
    if (m > mmax) LH_THROW(LH_ERR_ARG, "Chebyshev block width %d exceeds %d", m, mmax);
    if (half && m % 8 != 0) LH_THROW(LH_ERR_ARG, "FP16 Chebyshev storage needs a block width divisible by 8");
    if (!half && (m % 4 != 0 || vec_lanes(m, 4) < 4 || vec_lanes(m, 4) > 64))
        LH_THROW(LH_ERR_ARG, "time_step: block width %d not vectorisable", m);
    DevBuf w(n * m * 8), nrm(m * 8);
    randn<double>(ctx, n, m, w.as<double>(), m, 99);
    std::vector<double> hn(m, (double)n);  // unit-variance entries: ||w_c||^2 ~ n
    h2d(ctx, nrm.p, hn.data(), m * 8);
    void *xk, *xp;
    init(w.as<double>(), m, m, nrm.as<double>(), &xk, &xp);
    const double mu = (bmax - amin) / (bmax + amin);
    const double omega = 1.0 / (1.0 - 0.5 * mu * mu);
    hipEvent_t e0, e1;
    HIPX(hipEventCreate(&e0));
    HIPX(hipEventCreate(&e1));
    HIPX(hipEventRecord(e0, ctx->stream));
    for (int it = 0; it < iters; ++it) {
        step(m, R.p, xk, xp, omega);
        std::swap(xk, xp);
    }
    HIPX(hipEventRecord(e1, ctx->stream));
    HIPX(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPX(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 1000.0 * ms / iters;
}

}  // namespace lh

using namespace lh;

extern "C" int lh_time_cheb_step(lh_ctx* ctx, lh_csr* A, lh_csr* B, int ncol, int iters, int reorder, int storage,
                                 double* avg_us) {
    if (!ctx || !A || !B || !avg_us || iters <= 0 || ncol <= 0) return LH_ERR_ARG;
    if (storage != LH_STORE_F32 && storage != LH_STORE_F16) return LH_ERR_ARG;
    LH_API_BEGIN(ctx)
    HIPX(hipSetDevice(ctx->device));
    const lh_csr *a = A, *b = B;
    Reorder ro;
    std::unique_ptr<lh_csr> Ap, Bp;
    if (reorder && (B->is_diag || B->pat == A->pat)) {
        ro.build(ctx, *A->pat);
        Ap = ro.apply(ctx, A);
        Bp = ro.apply(ctx, B, B->is_diag ? nullptr : ro.pat);
        a = Ap.get();
        b = Bp.get();
    }
    ChebPrecond cp;
    cp.setup(ctx, a, b, -0.01, 2, 64.0, a->n(), ncol, storage == LH_STORE_F16);
    *avg_us = cp.time_step(ncol, iters);
    LH_API_END(ctx)
}
