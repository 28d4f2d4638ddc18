// Learned message-passing preconditioner (no counterpart in the reference,
// which factorises A - sigma B with SuperLU, lapy/solver.py:707-715).
//
// For the SPD operator K = beta A + alpha B (lumped mass mu = row sums of B)
// and residual columns R (n x c, row-major, leading dimension ld):
//
//     s = 1 / sqrt(diag K),  W^ = S (K - diag K) S       (FP32, on K's pattern)
//     phi_i = (1, log(d_i / mean d), log(mu_i / mean mu), sum_j W^_ij)
//     gam_l = 1 + tanh(phi Gam_l)                        (per vertex, F = 32)
//     Q x:   H0 = x u^T, A_l = W^ H_l,
//            H_{l+1} = gam_l * (H_l Ws_l^T + A_l Wn_l^T) (l = 0..2), Q x = H3 v
//     Q^T y: G3 = y v^T, T_l = gam_l * G_{l+1},
//            G_l = T_l Ws_l + W^ (T_l Wn_l),             Q^T y = G0 u
//     z = s (x + Q^T Q x), x = s r   P = S (I + Q^T Q) S: SPD; v = 0 is Jacobi
//
// Weights and hidden states are bfloat16 (states laid out [pair][feature],
// pair = vertex * columns + column, 64 bytes per pair); the layer products
// run on the matrix cores, v_mfma_f32_16x16x32_bf16 with FP32 accumulation;
// gathers, gates and sums are FP32.  A wave owns a tile of 16 pairs.  The
// products are formed transposed, D^T (32 output features x 16 pairs) =
// W (32 x 32) x H^T, so that the B operand is the tile's own rows: lane l
// loads 16 bytes (features 8q .. 8q+7, q = l / 16) of pair l % 16 and holds
// output features 4q .. 4q+3 of each 16-feature half of that pair.  The
// forward layer gathers the neighbours' rows of H_l into the aggregate A_l
// in the same 16-byte pieces, so the aggregation is fused with the product.
//
// Q^T runs on two kernels per layer (the product T_l Wn_l, then the local
// product and the neighbour gather).  Training: L = (1/c) sum_c
// ||S (K z_c - r_c)||^2 / ||S r_c||^2; with a = Q x and g_b = S dL/dz the
// gradient is the sum of two back-propagations through Q, (x, Q g_b) and
// (g_b, a), because <g_b, Q^T a> = <Q g_b, a>.  The weight gradients
// dWs = T^T H_l and dWn = T^T A_l also run on the matrix cores (the tile's
// rows staged in LDS and read transposed); every sum over pairs is taken per
// workgroup in a fixed order and over workgroups by a separate kernel in
// FP64, so the gradient is deterministic.
#include "lh_dense.hpp"

#include <cmath>
#include <cstring>
#include <random>

struct lh_mpnn {
    lh_ctx* ctx = nullptr;
    std::vector<float> params;  // host master copy
    DevBuf wbf;                 // device bfloat16 copy (+ transposed layer weights)
};

namespace lh {

namespace {

constexpr int F = 32;
constexpr int NL = 3;
constexpr int NFI = 4;  // per-vertex input features
constexpr int NPARAM = 2 * F + 2 * NL * F * F + NL * NFI * F;
constexpr int WB_LEN = NPARAM + 2 * NL * F * F;  // device weights: parameters, then Ws_l^T, Wn_l^T
constexpr int TBS = 256;                         // threads of the tile kernels: 4 waves, 64 pairs

// parameter offsets in the flat vector: u, (Ws_l, Wn_l) for l = 0..2, v, Gam_l
__host__ __device__ constexpr int off_u() { return 0; }
__host__ __device__ constexpr int off_ws(int l) { return F + l * 2 * F * F; }
__host__ __device__ constexpr int off_wn(int l) { return F + l * 2 * F * F + F * F; }
__host__ __device__ constexpr int off_v() { return F + NL * 2 * F * F; }
__host__ __device__ constexpr int off_gam(int l) { return 2 * F + NL * 2 * F * F + l * NFI * F; }
__host__ __device__ constexpr int off_wsT(int l) { return NPARAM + l * 2 * F * F; }
__host__ __device__ constexpr int off_wnT(int l) { return NPARAM + l * 2 * F * F + F * F; }

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__host__ __device__ __forceinline__ float bf2f(uint16_t b) {
#ifdef __HIP_DEVICE_COMPILE__
    return __uint_as_float((uint32_t)b << 16);
#else
    uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
#endif
}
// round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t f2bf(float f) {
    const uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
uint16_t f2bf_host(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ u16x8 ld8(const uint16_t* __restrict__ p) { return *reinterpret_cast<const u16x8*>(p); }
__device__ __forceinline__ bf16x8 as_bf(u16x8 v) { return __builtin_bit_cast(bf16x8, v); }
__device__ __forceinline__ f4 mma(u16x8 a, u16x8 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(a), as_bf(b), c, 0, 0, 0);
}
__device__ __forceinline__ u16x8 zero8() { return (u16x8){0, 0, 0, 0, 0, 0, 0, 0}; }

// the gates of every layer, vertex and feature, formed once per application:
// gam[l][i][o] = 1 + tanh(phi_i . Gam_l[:, o]) (FP32)
__global__ void k_mp_gates(int64_t n, const float4* __restrict__ phi, const uint16_t* __restrict__ wb,
                           float* __restrict__ gam) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= NL * n * F) return;
    const int o = (int)(gid % F);
    const int64_t li = gid / F;
    const int l = (int)(li / n);
    const int64_t i = li - (int64_t)l * n;
    const float4 ph = phi[i];
    const uint16_t* g = wb + off_gam(l) + o;
    const float z = ph.x * bf2f(g[0]) + ph.y * bf2f(g[F]) + ph.z * bf2f(g[2 * F]) + ph.w * bf2f(g[3 * F]);
    gam[gid] = 1.0f + tanhf(z);
}
__device__ __forceinline__ float4 ld4f(const float* __restrict__ p) { return *reinterpret_cast<const float4*>(p); }
__device__ __forceinline__ float f4at(const float4 v, int r) { return r == 0 ? v.x : r == 1 ? v.y : r == 2 ? v.z : v.w; }

// ---- set-up: s, W^ and the vertex features ---------------------------------
__global__ void k_mp_scale(int64_t n, const int32_t* __restrict__ diagpos, const double* __restrict__ kv,
                           float* __restrict__ s, double* __restrict__ dvec) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double d = diagpos[i] >= 0 ? kv[diagpos[i]] : 0.0;
    s[i] = d > 0.0 ? (float)(1.0 / sqrt(d)) : 0.0f;
    dvec[i] = d;
}
// w_slot = s_i K_ij s_j off the diagonal, 0 on it; wsum_i = sum of the row's w
__global__ void k_mp_weights(int64_t n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cols,
                             const double* __restrict__ kv, const float* __restrict__ s, float* __restrict__ w,
                             double* __restrict__ wsum) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double si = s[i];
    double t = 0.0;
    for (int32_t p = rowptr[i]; p < rowptr[i + 1]; ++p) {
        const int32_t j = cols[p];
        const float v = (j == (int32_t)i) ? 0.0f : (float)(si * kv[p] * (double)s[j]);
        w[p] = v;
        t += (double)v;
    }
    wsum[i] = t;
}
// lumped mass: row sums of B (on A's pattern or diagonal)
template <typename VT>
__global__ void k_mp_mass(int64_t n, const int32_t* __restrict__ rowptr, const VT* __restrict__ bv,
                          const VT* __restrict__ bdiag, double* __restrict__ mu) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double t = 0.0;
    if (bdiag) t = (double)bdiag[i];
    else
        for (int32_t p = rowptr[i]; p < rowptr[i + 1]; ++p) t += (double)bv[p];
    mu[i] = t;
}
// per-block partial sums and counts of the positive entries of a and b
__global__ void k_mp_possum(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                            double* __restrict__ part) {
    __shared__ double sh[4][256];
    double sa = 0, ca = 0, sb = 0, cb = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (a[i] > 0) { sa += a[i]; ca += 1; }
        if (b && b[i] > 0) { sb += b[i]; cb += 1; }
    }
    sh[0][threadIdx.x] = sa;
    sh[1][threadIdx.x] = ca;
    sh[2][threadIdx.x] = sb;
    sh[3][threadIdx.x] = cb;
    __syncthreads();
    if (threadIdx.x < 4) {
        double t = 0;
        for (int k = 0; k < 256; ++k) t += sh[threadIdx.x][k];
        part[blockIdx.x * 4 + threadIdx.x] = t;
    }
}
__global__ void k_mp_features(int64_t n, const double* __restrict__ d, const double* __restrict__ mu,
                              const double* __restrict__ wsum, const double* __restrict__ means,
                              float4* __restrict__ phi) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p;
    p.x = 1.0f;
    p.y = (d[i] > 0 && means[0] > 0) ? (float)log(d[i] / means[0]) : 0.0f;
    p.z = (mu && mu[i] > 0 && means[1] > 0) ? (float)log(mu[i] / means[1]) : 0.0f;
    p.w = (float)wsum[i];
    phi[i] = p;
}

// ---- lift and readout -------------------------------------------------------
// H[pc][f] = bf(x_pc w_f), x = X (leading dimension ldx) times s when given;
// one thread per (pair, 8 features)
template <typename XT>
__global__ void k_mp_lift(int64_t n, int nc, const XT* __restrict__ X, int ldx, const float* __restrict__ s,
                          const uint16_t* __restrict__ wb, int woff, uint16_t* __restrict__ H) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n * nc * 4) return;
    const int q = (int)(gid & 3);
    const int64_t pc = gid >> 2;
    const int64_t i = pc / nc;
    const int c = (int)(pc - i * nc);
    const float x = s ? (float)X[i * ldx + c] * s[i] : (float)X[i * ldx + c];
    u16x8 o;
#pragma unroll
    for (int j = 0; j < 8; ++j) o[j] = f2bf(x * bf2f(wb[woff + 8 * q + j]));
    *reinterpret_cast<u16x8*>(H + pc * F + 8 * q) = o;
}
// Y_pc = H[pc] . w (w = u or v)
__global__ void k_mp_readout(int64_t npair, const uint16_t* __restrict__ H, const uint16_t* __restrict__ wb, int woff,
                             double* __restrict__ Y) {
    const int64_t pc = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pc >= npair) return;
    float y = 0.f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const u16x8 h = ld8(H + pc * F + 8 * q);
#pragma unroll
        for (int j = 0; j < 8; ++j) y += bf2f(h[j]) * bf2f(wb[woff + 8 * q + j]);
    }
    Y[pc] = (double)y;
}
// z = s (s r + b), written over Z (may alias R: same element, same thread)
template <typename XT>
__global__ void k_mp_combine(int64_t n, int nc, const XT* R, int ldr, const float* __restrict__ s,
                             const double* __restrict__ Bv, XT* Z, int ldz) {
    const int64_t pc = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pc >= n * nc) return;
    const int64_t i = pc / nc;
    const int c = (int)(pc - i * nc);
    const float si = s[i];
    const float x = si * (float)R[i * ldr + c];
    Z[i * ldz + c] = (XT)(si * (x + (float)Bv[pc]));
}
// part[block][f] = sum over the block's pairs of y_pc H[pc][f] (y = X (ldx)
// times s when given): dv (H = H3, y = dy) and du (H = dH0, y = x)
__global__ void __launch_bounds__(256)
k_mp_outer(int64_t n, int nc, const double* __restrict__ X, int ldx, const float* __restrict__ s,
           const uint16_t* __restrict__ H, double* __restrict__ part) {
    __shared__ float red[64][F + 1];
    const int q = threadIdx.x & 3;
    const int64_t npair = n * nc;
    float acc[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[j] = 0.f;
    for (int64_t pc = (int64_t)blockIdx.x * 64 + (threadIdx.x >> 2); pc < npair; pc += (int64_t)gridDim.x * 64) {
        const int64_t i = pc / nc;
        const int c = (int)(pc - i * nc);
        const float y = s ? (float)X[i * ldx + c] * s[i] : (float)X[i * ldx + c];
        const u16x8 h = ld8(H + pc * F + 8 * q);
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[j] += y * bf2f(h[j]);
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) red[threadIdx.x >> 2][8 * q + j] = acc[j];
    __syncthreads();
    if (threadIdx.x < F) {
        double t = 0.0;
        for (int r = 0; r < 64; ++r) t += (double)red[r][threadIdx.x];
        part[(int64_t)blockIdx.x * F + threadIdx.x] = t;
    }
}

// ---- forward layer -------------------------------------------------------------
// Hout = bf(gam_l * (Hin Ws^T + A Wn^T)), A = bf(W^ Hin) (stored in Aout
// when given).  Tiles of 16 pairs, one per wave, grid-stride.
__global__ void __launch_bounds__(TBS)
k_mp_fwd(int64_t n, int nc, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cols,
         const float* __restrict__ w, const float* __restrict__ gam, const uint16_t* __restrict__ Hin,
         const uint16_t* __restrict__ wb, int l, uint16_t* __restrict__ Hout, uint16_t* __restrict__ Aout) {
    const int lane = threadIdx.x & 63;
    const int pn = lane & 15, q = lane >> 4;
    const int64_t npair = n * nc;
    const int64_t ntile = (npair + 15) / 16;
    // A operands: rows o = hf*16 + pn of Ws and Wn, features 8q .. 8q+7
    u16x8 aws[2], awn[2];
#pragma unroll
    for (int hf = 0; hf < 2; ++hf) {
        aws[hf] = ld8(wb + off_ws(l) + (hf * 16 + pn) * F + 8 * q);
        awn[hf] = ld8(wb + off_wn(l) + (hf * 16 + pn) * F + 8 * q);
    }
    // each workgroup walks a contiguous range of tiles, the ranges of one XCD
    // adjacent (gridDim.x is a multiple of 8): in the solver's locality
    // order the gathered rows then sit in that XCD's L2
    const int64_t tpb = (ntile + gridDim.x - 1) / gridDim.x;
    const int64_t tb0 = xcd_block() * tpb, tb1 = tb0 + tpb < ntile ? tb0 + tpb : ntile;
    for (int64_t t = tb0 + (threadIdx.x >> 6); t < tb1; t += TBS / 64) {
        const int64_t p = t * 16 + pn;
        const bool ok = p < npair;
        u16x8 h = zero8(), a = zero8();
        int64_t i = 0;
        if (ok) {
            i = p / nc;
            const int c = (int)(p - i * nc);
            h = ld8(Hin + p * F + 8 * q);
            float g[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) g[j] = 0.f;
            // the row's slots eight at a time (a surface row has at most eight),
            // all their column indices, then all their rows, loaded together;
            // slots past the row's end and the diagonal slot have weight 0
            const int32_t e0 = rowptr[i], e1 = rowptr[i + 1];
            for (int32_t e = e0; e < e1; e += 8) {
                int32_t cj[8];
                float wj[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const bool in = e + u < e1;
                    cj[u] = in ? cols[e + u] : (int32_t)i;
                    wj[u] = in ? w[e + u] : 0.f;
                }
                u16x8 hj[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) hj[u] = ld8(Hin + ((int64_t)cj[u] * nc + c) * F + 8 * q);
#pragma unroll
                for (int u = 0; u < 8; ++u)
#pragma unroll
                    for (int j = 0; j < 8; ++j) g[j] += wj[u] * bf2f(hj[u][j]);
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) a[j] = f2bf(g[j]);
            if (Aout) *reinterpret_cast<u16x8*>(Aout + p * F + 8 * q) = a;
        }
        f4 d[2];
#pragma unroll
        for (int hf = 0; hf < 2; ++hf) {
            d[hf] = (f4){0.f, 0.f, 0.f, 0.f};
            d[hf] = mma(aws[hf], h, d[hf]);
            d[hf] = mma(awn[hf], a, d[hf]);
        }
        if (!ok) continue;
        const float* gi = gam + ((int64_t)l * n + i) * F;
#pragma unroll
        for (int hf = 0; hf < 2; ++hf) {
            const float4 g4 = ld4f(gi + hf * 16 + 4 * q);
            u16x4 o;
#pragma unroll
            for (int r = 0; r < 4; ++r) o[r] = f2bf(f4at(g4, r) * d[hf][r]);
            *reinterpret_cast<u16x4*>(Hout + p * F + hf * 16 + 4 * q) = o;
        }
    }
}

// ---- adjoint / back-propagation -------------------------------------------------
// Part 1 of a layer: T = bf(gam_l * G) (G = G_{l+1} or dH_{l+1}), M = bf(T Wn)
// into Mout.  With grads (part != null): also the weight-gradient partials
// dWs += T^T Hl, dWn += T^T Al and dGam += phi^T (G * pre * (1 - tanh^2)),
// pre = Hl Ws^T + Al Wn^T, per workgroup into part[block][2 F F + NFI F].
// LDS of the weight-gradient form (a template so that the inference form has none)
template <bool GRADS>
struct BwdShared;
template <>
struct BwdShared<false> {
    uint16_t sT[1][8], sH[1][8], sA[1][8];  // not used by the inference form
};
template <>
struct BwdShared<true> {
    uint16_t sT[4][16 * F], sH[4][16 * F], sA[4][16 * F];
    float red[4][64][40];
};
template <bool GRADS>
__global__ void __launch_bounds__(TBS)
k_mp_bwd_m(int64_t n, int nc, const float4* __restrict__ phi, const float* __restrict__ gam,
           const uint16_t* __restrict__ G, const uint16_t* __restrict__ Hl, const uint16_t* __restrict__ Al,
           const uint16_t* __restrict__ wb, int l, uint16_t* __restrict__ Mout, double* __restrict__ part) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int pn = lane & 15, q = lane >> 4;
    const int64_t npair = n * nc;
    const int64_t ntile = (npair + 15) / 16;
    constexpr bool grads = GRADS;
    __shared__ BwdShared<GRADS> shm;
    auto& sT = shm.sT;
    auto& sH = shm.sH;
    auto& sA = shm.sA;
    u16x8 awnT[2], aws[2], awn[2];
#pragma unroll
    for (int hf = 0; hf < 2; ++hf) {
        awnT[hf] = ld8(wb + off_wnT(l) + (hf * 16 + pn) * F + 8 * q);
        aws[hf] = ld8(wb + off_ws(l) + (hf * 16 + pn) * F + 8 * q);
        awn[hf] = ld8(wb + off_wn(l) + (hf * 16 + pn) * F + 8 * q);
    }
    f4 gws[2][2], gwn[2][2];  // [ob][kb]: dW[ob*16 + 4q + r][kb*16 + pn]
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) gws[a][b] = gwn[a][b] = (f4){0.f, 0.f, 0.f, 0.f};
    float ggam[NFI][8];  // dGam[k][hf*16 + 4q + r], index hf*4 + r
#pragma unroll
    for (int k = 0; k < NFI; ++k)
#pragma unroll
        for (int j = 0; j < 8; ++j) ggam[k][j] = 0.f;
    const int64_t tpb = (ntile + gridDim.x - 1) / gridDim.x;
    const int64_t tb0 = xcd_block() * tpb, tb1 = tb0 + tpb < ntile ? tb0 + tpb : ntile;
    for (int64_t t = tb0 + wv; t < tb1; t += TBS / 64) {
        const int64_t p = t * 16 + pn;
        const bool ok = p < npair;
        u16x8 tv = zero8();
        float4 ph = make_float4(0.f, 0.f, 0.f, 0.f);
        const float* gi = gam;
        if (ok) {
            const int64_t i = p / nc;
            if (GRADS) ph = phi[i];
            gi = gam + ((int64_t)l * n + i) * F;
            const u16x8 gv = ld8(G + p * F + 8 * q);
            const float4 ga = ld4f(gi + 8 * q), gb = ld4f(gi + 8 * q + 4);
#pragma unroll
            for (int j = 0; j < 8; ++j) tv[j] = f2bf((j < 4 ? f4at(ga, j) : f4at(gb, j - 4)) * bf2f(gv[j]));
        }
        f4 d[2];
#pragma unroll
        for (int hf = 0; hf < 2; ++hf) d[hf] = mma(awnT[hf], tv, (f4){0.f, 0.f, 0.f, 0.f});
        if (ok)
#pragma unroll
            for (int hf = 0; hf < 2; ++hf) {
                u16x4 o;
#pragma unroll
                for (int r = 0; r < 4; ++r) o[r] = f2bf(d[hf][r]);
                *reinterpret_cast<u16x4*>(Mout + p * F + hf * 16 + 4 * q) = o;
            }
        if (!grads) continue;
        // pre (lane: pair pn, features hf*16 + 4q + r) and the gate gradient
        u16x8 hv = zero8(), av = zero8();
        if (ok) {
            hv = ld8(Hl + p * F + 8 * q);
            av = ld8(Al + p * F + 8 * q);
        }
        f4 pre[2];
#pragma unroll
        for (int hf = 0; hf < 2; ++hf) {
            pre[hf] = mma(aws[hf], hv, (f4){0.f, 0.f, 0.f, 0.f});
            pre[hf] = mma(awn[hf], av, pre[hf]);
        }
        if (ok)
#pragma unroll
            for (int hf = 0; hf < 2; ++hf) {
                const u16x4 g4 = *reinterpret_cast<const u16x4*>(G + p * F + hf * 16 + 4 * q);
                const float4 gm = ld4f(gi + hf * 16 + 4 * q);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const float th = f4at(gm, r) - 1.0f;  // tanh(z)
                    const float dg = bf2f(g4[r]) * pre[hf][r] * (1.0f - th * th);
                    ggam[0][hf * 4 + r] += ph.x * dg;
                    ggam[1][hf * 4 + r] += ph.y * dg;
                    ggam[2][hf * 4 + r] += ph.z * dg;
                    ggam[3][hf * 4 + r] += ph.w * dg;
                }
            }
        // stage the tile's T, Hl, Al rows ([pair][feature]) for the transposed reads
        *reinterpret_cast<u16x8*>(&sT[wv][pn * F + 8 * q]) = tv;
        *reinterpret_cast<u16x8*>(&sH[wv][pn * F + 8 * q]) = hv;
        *reinterpret_cast<u16x8*>(&sA[wv][pn * F + 8 * q]) = av;
        __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): this wave's LDS stores are done
        __builtin_amdgcn_wave_barrier();
        // dW[o][k] += sum_pairs T[pair][o] X[pair][k]: A[m = o][kk = pair],
        // B[kk = pair][n = k]; pairs 16..31 of the 32-deep product are zero
        u16x8 at[2], bh[2], ba[2];
#pragma unroll
        for (int ob = 0; ob < 2; ++ob) {
            at[ob] = zero8();
            bh[ob] = zero8();
            ba[ob] = zero8();
            if (q < 2)
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int pr = 8 * q + j;
                    at[ob][j] = sT[wv][pr * F + ob * 16 + pn];
                    bh[ob][j] = sH[wv][pr * F + ob * 16 + pn];
                    ba[ob][j] = sA[wv][pr * F + ob * 16 + pn];
                }
        }
        __builtin_amdgcn_wave_barrier();  // reads done before the next tile's stores
#pragma unroll
        for (int ob = 0; ob < 2; ++ob)
#pragma unroll
            for (int kb = 0; kb < 2; ++kb) {
                gws[ob][kb] = mma(at[ob], bh[kb], gws[ob][kb]);
                gwn[ob][kb] = mma(at[ob], ba[kb], gwn[ob][kb]);
            }
    }
    if constexpr (!GRADS) return;
    else {
    auto& red = shm.red;
    // workgroup sums in a fixed order: each wave's lanes write their values,
    // then one thread per gradient entry adds them up
    float* mine = red[wv][lane];
#pragma unroll
    for (int ob = 0; ob < 2; ++ob)
#pragma unroll
        for (int kb = 0; kb < 2; ++kb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                mine[(ob * 2 + kb) * 4 + r] = gws[ob][kb][r];
                mine[16 + (ob * 2 + kb) * 4 + r] = gwn[ob][kb][r];
            }
    __syncthreads();
    double* out = part + (int64_t)blockIdx.x * (2 * F * F + NFI * F);
    for (int e = threadIdx.x; e < 2 * F * F; e += TBS) {
        const int which = e / (F * F);  // 0: Ws, 1: Wn
        const int o = (e % (F * F)) / F, k = e % F;
        const int ob = o / 16, rq = o % 16, kb = k / 16, pl = k % 16;
        const int qq = rq / 4, r = rq % 4;
        const int ln = qq * 16 + pl;
        double tsum = 0.0;
        for (int w2 = 0; w2 < 4; ++w2) tsum += (double)red[w2][ln][which * 16 + (ob * 2 + kb) * 4 + r];
        out[e] = tsum;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NFI; ++k)
#pragma unroll
        for (int j = 0; j < 8; ++j) mine[k * 8 + j] = ggam[k][j];
    __syncthreads();
    for (int e = threadIdx.x; e < NFI * F; e += TBS) {
        const int k = e / F, o = e % F;
        const int hf = o / 16, qq = (o % 16) / 4, r = o % 4;
        double tsum = 0.0;
        for (int w2 = 0; w2 < 4; ++w2)
            for (int pl = 0; pl < 16; ++pl) tsum += (double)red[w2][qq * 16 + pl][k * 8 + hf * 4 + r];
        out[2 * F * F + e] = tsum;
    }
    }
}
// Part 2: Gout = bf(T Ws + W^ M), T = bf(gam_l * G) recomputed
__global__ void __launch_bounds__(TBS)
k_mp_bwd_g(int64_t n, int nc, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cols,
           const float* __restrict__ w, const float* __restrict__ gam, const uint16_t* __restrict__ G,
           const uint16_t* __restrict__ M, const uint16_t* __restrict__ wb, int l, uint16_t* __restrict__ Gout) {
    const int lane = threadIdx.x & 63;
    const int pn = lane & 15, q = lane >> 4;
    const int64_t npair = n * nc;
    const int64_t ntile = (npair + 15) / 16;
    u16x8 awsT[2];
#pragma unroll
    for (int hf = 0; hf < 2; ++hf) awsT[hf] = ld8(wb + off_wsT(l) + (hf * 16 + pn) * F + 8 * q);
    // each workgroup walks a contiguous range of tiles, the ranges of one XCD
    // adjacent (gridDim.x is a multiple of 8): in the solver's locality
    // order the gathered rows then sit in that XCD's L2
    const int64_t tpb = (ntile + gridDim.x - 1) / gridDim.x;
    const int64_t tb0 = xcd_block() * tpb, tb1 = tb0 + tpb < ntile ? tb0 + tpb : ntile;
    for (int64_t t = tb0 + (threadIdx.x >> 6); t < tb1; t += TBS / 64) {
        const int64_t p = t * 16 + pn;
        const bool ok = p < npair;
        u16x8 tv = zero8();
        if (ok) {
            const float* gi = gam + ((int64_t)l * n + p / nc) * F;
            const u16x8 gv = ld8(G + p * F + 8 * q);
            const float4 ga = ld4f(gi + 8 * q), gb = ld4f(gi + 8 * q + 4);
#pragma unroll
            for (int j = 0; j < 8; ++j) tv[j] = f2bf((j < 4 ? f4at(ga, j) : f4at(gb, j - 4)) * bf2f(gv[j]));
        }
        f4 d[2];
#pragma unroll
        for (int hf = 0; hf < 2; ++hf) d[hf] = mma(awsT[hf], tv, (f4){0.f, 0.f, 0.f, 0.f});
        if (!ok) continue;
        const int64_t i = p / nc;
        const int c = (int)(p - i * nc);
        float g[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) g[j] = 0.f;
        // eight slots at a time, indices then rows loaded together (weight 0
        // past the row's end and on the diagonal)
        const int32_t e0 = rowptr[i], e1 = rowptr[i + 1];
        for (int32_t e = e0; e < e1; e += 8) {
            int32_t cj[8];
            float wj[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const bool in = e + u < e1;
                cj[u] = in ? cols[e + u] : (int32_t)i;
                wj[u] = in ? w[e + u] : 0.f;
            }
            u16x4 m0[8], m1[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const uint16_t* mj = M + ((int64_t)cj[u] * nc + c) * F + 4 * q;
                m0[u] = *reinterpret_cast<const u16x4*>(mj);
                m1[u] = *reinterpret_cast<const u16x4*>(mj + 16);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    g[r] += wj[u] * bf2f(m0[u][r]);
                    g[4 + r] += wj[u] * bf2f(m1[u][r]);
                }
        }
#pragma unroll
        for (int hf = 0; hf < 2; ++hf) {
            u16x4 o;
#pragma unroll
            for (int r = 0; r < 4; ++r) o[r] = f2bf(d[hf][r] + g[hf * 4 + r]);
            *reinterpret_cast<u16x4*>(Gout + p * F + hf * 16 + 4 * q) = o;
        }
    }
}

// ---- training -----------------------------------------------------------
// e = s (KZ - R) (into KZ), t = s e (into T), s r (into SR)
__global__ void k_mp_err(int64_t n, int nc, double* __restrict__ KZ, const double* __restrict__ R,
                         const float* __restrict__ s, double* __restrict__ T, double* __restrict__ SR) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n * nc) return;
    const int64_t i = gid / nc;
    const double si = s[i];
    const double e = si * (KZ[gid] - R[gid]);
    KZ[gid] = e;
    T[gid] = si * e;
    SR[gid] = si * R[gid];
}
__global__ void k_mp_rowscale(int64_t n, int nc, const float* __restrict__ s, double* __restrict__ X) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n * nc) return;
    X[gid] *= (double)s[gid / nc];
}

// out[j] = sum over workgroups of part[b][j] (fixed order)
__global__ void k_mp_reduce(int nb, int len, int stride, const double* __restrict__ part, double* __restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= len) return;
    double t = 0.0;
    for (int b = 0; b < nb; ++b) t += part[(int64_t)b * stride + j];
    out[j] = t;
}

constexpr int kPairBlocks = 1024;  // workgroups of the grid-stride pair loops

}  // namespace

// The operator-dependent state (K, s, W^, phi) and the work buffers of the
// network; reused across applications with the same K.
struct MpnnRun {
    lh_ctx* ctx = nullptr;
    std::shared_ptr<Pattern> pat;
    int64_t n = 0;
    DevBuf kv, s, w, phi, gam;
    DevBuf HA[NL + 1], AA[NL], HB[NL + 1], AB[NL];  // states of two forward passes (bfloat16)
    DevBuf G0, G1, Mb;                              // adjoint / back-propagation (bfloat16)
    DevBuf fa, fb;                                  // scalar fields (n x nc, FP64)

    void setup(lh_ctx* c, const lh_csr* A, const lh_csr* B, double alpha, double beta) {
        ctx = c;
        pat = A->pat;
        n = A->n();
        form_shifted(ctx, A, B, alpha, beta, kv);
        s.ensure(n * 4);
        w.ensure(pat->nnz * 4);
        phi.ensure(n * 16);
        DevBuf dvec(n * 8), wsum(n * 8), mu, part((size_t)256 * 4 * 8), means(16);
        k_mp_scale<<<nblocks(n, 256), 256, 0, ctx->stream>>>(n, pat->diagpos.as<int32_t>(), kv.as<double>(),
                                                            s.as<float>(), dvec.as<double>());
        k_mp_weights<<<nblocks(n, 256), 256, 0, ctx->stream>>>(n, pat->rowptr.as<int32_t>(), pat->cols.as<int32_t>(),
                                                              kv.as<double>(), s.as<float>(), w.as<float>(),
                                                              wsum.as<double>());
        LH_CHECK_LAUNCH();
        const bool has_mass = B != nullptr && alpha != 0.0;
        if (has_mass) {
            mu.alloc(n * 8);
            if (B->dtype == LH_F64)
                k_mp_mass<double><<<nblocks(n, 256), 256, 0, ctx->stream>>>(
                    n, B->pat->rowptr.as<int32_t>(), B->vals.as<double>(),
                    B->is_diag ? B->dense_diag.as<double>() : nullptr, mu.as<double>());
            else
                k_mp_mass<float><<<nblocks(n, 256), 256, 0, ctx->stream>>>(
                    n, B->pat->rowptr.as<int32_t>(), B->vals.as<float>(),
                    B->is_diag ? B->dense_diag.as<float>() : nullptr, mu.as<double>());
            LH_CHECK_LAUNCH();
        }
        k_mp_possum<<<256, 256, 0, ctx->stream>>>(n, dvec.as<double>(), has_mass ? mu.as<double>() : nullptr,
                                                  part.as<double>());
        LH_CHECK_LAUNCH();
        std::vector<double> hp(256 * 4);
        d2h(ctx, hp.data(), part.p, hp.size() * 8);
        double sd = 0, cd = 0, sm = 0, cm = 0;
        for (int b = 0; b < 256; ++b) {
            sd += hp[b * 4];
            cd += hp[b * 4 + 1];
            sm += hp[b * 4 + 2];
            cm += hp[b * 4 + 3];
        }
        const double hm[2] = {cd > 0 ? sd / cd : 0.0, cm > 0 ? sm / cm : 0.0};
        h2d(ctx, means.p, hm, 16);
        k_mp_features<<<nblocks(n, 256), 256, 0, ctx->stream>>>(n, dvec.as<double>(),
                                                               has_mass ? mu.as<double>() : nullptr,
                                                               wsum.as<double>(), means.as<double>(),
                                                               phi.as<float4>());
        LH_CHECK_LAUNCH();
        HIPX(hipStreamSynchronize(ctx->stream));  // the temporaries go out of scope
    }
    // the gates of the current weights (before each application)
    void gates(const uint16_t* wb) {
        gam.ensure((size_t)NL * n * F * 4);
        k_mp_gates<<<nblocks((int64_t)NL * n * F, 256), 256, 0, ctx->stream>>>(n, phi.as<float4>(), wb,
                                                                              gam.as<float>());
        LH_CHECK_LAUNCH();
    }
    size_t hbytes(int nc) const { return (size_t)n * nc * F * 2; }
    // workgroups of the tile kernels: a multiple of 8 (one range per XCD slot)
    unsigned tile_blocks(int nc) const {
        const int64_t ntile = (n * nc + 15) / 16;
        const int64_t nb = std::max<int64_t>(1, std::min<int64_t>((ntile + 3) / 4, (int64_t)ctx->n_cu * 8));
        return (unsigned)((nb + kXCDs - 1) / kXCDs * kXCDs);
    }
    // Y (dense n x nc) = Q x, x = X (ldx), times s when scaled.  store: keep
    // every layer's input and aggregate in H, A; else ping-pong H[0], H[1]
    template <typename XT>
    void q_forward(const uint16_t* wb, const XT* X, int ldx, bool scaled, int nc, DevBuf* H, DevBuf* A,
                   bool store, double* Y) {
        for (int l = 0; l <= (store ? NL : 1); ++l) H[l].ensure(hbytes(nc));
        if (store)
            for (int l = 0; l < NL; ++l) A[l].ensure(hbytes(nc));
        k_mp_lift<<<nblocks(n * nc * 4, 256), 256, 0, ctx->stream>>>(n, nc, X, ldx, scaled ? s.as<float>() : nullptr,
                                                                    wb, off_u(), H[0].as<uint16_t>());
        for (int l = 0; l < NL; ++l) {
            const uint16_t* hin = H[store ? l : (l & 1)].as<uint16_t>();
            uint16_t* hout = H[store ? l + 1 : ((l + 1) & 1)].as<uint16_t>();
            k_mp_fwd<<<tile_blocks(nc), TBS, 0, ctx->stream>>>(n, nc, pat->rowptr.as<int32_t>(),
                                                              pat->cols.as<int32_t>(), w.as<float>(),
                                                              gam.as<float>(), hin, wb, l, hout,
                                                              store ? A[l].as<uint16_t>() : nullptr);
        }
        k_mp_readout<<<nblocks(n * nc, 256), 256, 0, ctx->stream>>>(n * nc, H[store ? NL : (NL & 1)].as<uint16_t>(),
                                                                   wb, off_v(), Y);
        LH_CHECK_LAUNCH();
    }
    // out (dense n x nc) = Q^T Y
    void q_adjoint(const uint16_t* wb, const double* Y, int nc, double* out) {
        G0.ensure(hbytes(nc));
        G1.ensure(hbytes(nc));
        Mb.ensure(hbytes(nc));
        k_mp_lift<<<nblocks(n * nc * 4, 256), 256, 0, ctx->stream>>>(n, nc, Y, nc, nullptr, wb, off_v(),
                                                                    G0.as<uint16_t>());
        DevBuf* g = &G0;
        DevBuf* gn = &G1;
        for (int l = NL - 1; l >= 0; --l) {
            k_mp_bwd_m<false><<<tile_blocks(nc), TBS, 0, ctx->stream>>>(n, nc, phi.as<float4>(), gam.as<float>(),
                                                                       g->as<uint16_t>(), nullptr, nullptr, wb, l,
                                                                       Mb.as<uint16_t>(), nullptr);
            k_mp_bwd_g<<<tile_blocks(nc), TBS, 0, ctx->stream>>>(n, nc, pat->rowptr.as<int32_t>(),
                                                                pat->cols.as<int32_t>(), w.as<float>(),
                                                                gam.as<float>(), g->as<uint16_t>(),
                                                                Mb.as<uint16_t>(), wb, l, gn->as<uint16_t>());
            std::swap(g, gn);
        }
        k_mp_readout<<<nblocks(n * nc, 256), 256, 0, ctx->stream>>>(n * nc, g->as<uint16_t>(), wb, off_u(), out);
        LH_CHECK_LAUNCH();
    }
    // gdev (NPARAM doubles) = gradient of <dY, Q x> over the parameters, with
    // the stored states H, A of input x = X (ldx) (times s when scaled)
    void q_backprop(const uint16_t* wb, const double* X, int ldx, bool scaled, int nc, DevBuf* H, DevBuf* A,
                    const double* dY, DevBuf& partv, DevBuf& partw, double* gdev) {
        G0.ensure(hbytes(nc));
        G1.ensure(hbytes(nc));
        Mb.ensure(hbytes(nc));
        constexpr int WLEN = 2 * F * F + NFI * F;
        // dv = H3^T dY, dH3 = bf(dY v^T)
        k_mp_outer<<<kPairBlocks, 256, 0, ctx->stream>>>(n, nc, dY, nc, nullptr, H[NL].as<uint16_t>(),
                                                         partv.as<double>());
        k_mp_reduce<<<1, 64, 0, ctx->stream>>>(kPairBlocks, F, F, partv.as<double>(), gdev + off_v());
        k_mp_lift<<<nblocks(n * nc * 4, 256), 256, 0, ctx->stream>>>(n, nc, dY, nc, nullptr, wb, off_v(),
                                                                    G0.as<uint16_t>());
        DevBuf* g = &G0;
        DevBuf* gn = &G1;
        const unsigned tb = tile_blocks(nc);
        for (int l = NL - 1; l >= 0; --l) {
            k_mp_bwd_m<true><<<tb, TBS, 0, ctx->stream>>>(n, nc, phi.as<float4>(), gam.as<float>(),
                                                          g->as<uint16_t>(), H[l].as<uint16_t>(),
                                                          A[l].as<uint16_t>(), wb, l, Mb.as<uint16_t>(),
                                                          partw.as<double>());
            k_mp_reduce<<<nblocks(2 * F * F, 256), 256, 0, ctx->stream>>>((int)tb, 2 * F * F, WLEN,
                                                                          partw.as<double>(), gdev + off_ws(l));
            k_mp_reduce<<<1, NFI * F, 0, ctx->stream>>>((int)tb, NFI * F, WLEN, partw.as<double>() + 2 * F * F,
                                                        gdev + off_gam(l));
            k_mp_bwd_g<<<tb, TBS, 0, ctx->stream>>>(n, nc, pat->rowptr.as<int32_t>(), pat->cols.as<int32_t>(),
                                                    w.as<float>(), gam.as<float>(), g->as<uint16_t>(),
                                                    Mb.as<uint16_t>(), wb, l, gn->as<uint16_t>());
            std::swap(g, gn);
        }
        // du = dH0^T x
        k_mp_outer<<<kPairBlocks, 256, 0, ctx->stream>>>(n, nc, X, ldx, scaled ? s.as<float>() : nullptr,
                                                         g->as<uint16_t>(), partv.as<double>());
        k_mp_reduce<<<1, 64, 0, ctx->stream>>>(kPairBlocks, F, F, partv.as<double>(), gdev + off_u());
        LH_CHECK_LAUNCH();
    }
    // W (n x m, ldw) <- P W = S (I + Q^T Q) S W, in place (FP64 or FP32 blocks)
    template <typename XT>
    void apply(const uint16_t* wb, XT* W, int ldw, int m) {
        fa.ensure((size_t)n * m * 8);
        fb.ensure((size_t)n * m * 8);
        gates(wb);
        q_forward(wb, W, ldw, true, m, HA, AA, false, fa.as<double>());
        q_adjoint(wb, fa.as<double>(), m, fb.as<double>());
        k_mp_combine<<<nblocks(n * m, 256), 256, 0, ctx->stream>>>(n, m, W, ldw, s.as<float>(), fb.as<double>(), W,
                                                                  ldw);
        LH_CHECK_LAUNCH();
    }
    // average duration (us) of the three forward layers on an n x m block
    double time_layers(const uint16_t* wb, int m, int iters) {
        for (int l = 0; l < 2; ++l) HA[l].ensure(hbytes(m));
        AA[0].ensure(hbytes(m));
        HIPX(hipMemsetAsync(HA[0].p, 0, hbytes(m), ctx->stream));
        gates(wb);
        hipEvent_t e0, e1;
        HIPX(hipEventCreate(&e0));
        HIPX(hipEventCreate(&e1));
        HIPX(hipEventRecord(e0, ctx->stream));
        for (int it = 0; it < iters; ++it)
            k_mp_fwd<<<tile_blocks(m), TBS, 0, ctx->stream>>>(n, m, pat->rowptr.as<int32_t>(), pat->cols.as<int32_t>(),
                                                             w.as<float>(), gam.as<float>(), HA[0].as<uint16_t>(),
                                                             wb, it % NL, HA[1].as<uint16_t>(), nullptr);
        HIPX(hipEventRecord(e1, ctx->stream));
        HIPX(hipEventSynchronize(e1));
        LH_CHECK_LAUNCH();
        float ms = 0.f;
        HIPX(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return 1000.0 * ms / iters;
    }
};

// loss and gradient (host) for residual columns R (device, n x nc, dense)
void mpnn_loss_grad(MpnnRun& run, const uint16_t* wb, const double* R, int nc, double* loss, double* grad) {
    lh_ctx* ctx = run.ctx;
    const int64_t n = run.n;
    DevBuf a(n * nc * 8), b(n * nc * 8), Z(n * nc * 8), KZ(n * nc * 8), Tt(n * nc * 8), SR(n * nc * 8),
        gb(n * nc * 8), ga(n * nc * 8);
    // z = s (x + Q^T Q x), x = s r
    run.gates(wb);
    run.q_forward(wb, R, nc, true, nc, run.HA, run.AA, true, a.as<double>());
    run.q_adjoint(wb, a.as<double>(), nc, b.as<double>());
    k_mp_combine<<<nblocks(n * nc, 256), 256, 0, ctx->stream>>>(n, nc, R, nc, run.s.as<float>(), b.as<double>(),
                                                               Z.as<double>(), nc);
    LH_CHECK_LAUNCH();
    lh_csr Kv;
    Kv.ctx = ctx;
    Kv.pat = run.pat;
    Kv.dtype = LH_F64;
    Kv.vals.p = run.kv.p;
    Kv.vals.pooled = true;  // borrowed: not freed by this view
    spmm_ld<double, double>(ctx, &Kv, Z.as<double>(), KZ.as<double>(), nc, nc, nc);
    k_mp_err<<<nblocks(n * nc, 256), 256, 0, ctx->stream>>>(n, nc, KZ.as<double>(), R, run.s.as<float>(),
                                                           Tt.as<double>(), SR.as<double>());
    LH_CHECK_LAUNCH();
    DevBuf cws, e2(nc * 8), r2(nc * 8);
    colsq<double>(ctx, n, nc, KZ.as<double>(), nc, e2.as<double>(), cws);
    colsq<double>(ctx, n, nc, SR.as<double>(), nc, r2.as<double>(), cws);
    std::vector<double> he(nc), hr(nc), scl(nc);
    d2h(ctx, he.data(), e2.p, nc * 8);
    d2h(ctx, hr.data(), r2.p, nc * 8);
    double L = 0.0;
    for (int c = 0; c < nc; ++c) {
        L += he[c] / hr[c] / nc;
        scl[c] = 2.0 / (hr[c] * nc);
    }
    *loss = L;
    // g_b = s dL/dz = s K (s e) * 2 / (nc ||s r_c||^2)
    spmm_ld<double, double>(ctx, &Kv, Tt.as<double>(), gb.as<double>(), nc, nc, nc);
    Kv.vals.p = nullptr;
    DevBuf dscl(nc * 8);
    h2d(ctx, dscl.p, scl.data(), nc * 8);
    colscale<double>(ctx, n, nc, gb.as<double>(), nc, dscl.as<double>());
    k_mp_rowscale<<<nblocks(n * nc, 256), 256, 0, ctx->stream>>>(n, nc, run.s.as<float>(), gb.as<double>());
    LH_CHECK_LAUNCH();
    // g_a = Q g_b (states kept for the second back-propagation)
    run.q_forward(wb, gb.as<double>(), nc, false, nc, run.HB, run.AB, true, ga.as<double>());
    // dL/dtheta = grad <g_a, Q x> + grad <a, Q g_b>
    const unsigned tb = run.tile_blocks(nc);
    DevBuf partv((size_t)kPairBlocks * F * 8), partw((size_t)tb * (2 * F * F + NFI * F) * 8), g1(NPARAM * 8),
        g2(NPARAM * 8);
    run.q_backprop(wb, R, nc, true, nc, run.HA, run.AA, ga.as<double>(), partv, partw, g1.as<double>());
    run.q_backprop(wb, gb.as<double>(), nc, false, nc, run.HB, run.AB, a.as<double>(), partv, partw, g2.as<double>());
    std::vector<double> h1(NPARAM), h2(NPARAM);
    d2h(ctx, h1.data(), g1.p, NPARAM * 8);
    d2h(ctx, h2.data(), g2.p, NPARAM * 8);
    for (int i = 0; i < NPARAM; ++i) grad[i] = h1[i] + h2[i];
}

int mpnn_param_count() { return NPARAM; }

void mpnn_upload(lh_mpnn* net) {
    std::vector<uint16_t> b(WB_LEN);
    for (int i = 0; i < NPARAM; ++i) b[i] = f2bf_host(net->params[i]);
    for (int l = 0; l < NL; ++l)
        for (int o = 0; o < F; ++o)
            for (int k = 0; k < F; ++k) {
                b[off_wsT(l) + o * F + k] = b[off_ws(l) + k * F + o];
                b[off_wnT(l) + o * F + k] = b[off_wn(l) + k * F + o];
            }
    net->wbf.persistent = true;
    net->wbf.ensure(WB_LEN * 2);
    h2d(net->ctx, net->wbf.p, b.data(), WB_LEN * 2);
}

}  // namespace lh

using namespace lh;

// LearnedPrecond (declared in lh_dense.hpp): the eigensolver's / PCG's handle
void LearnedPrecond::setup(lh_ctx* c, lh_mpnn* net_, const lh_csr* A, const lh_csr* B, double alpha, double beta) {
    net = net_;
    run = std::make_unique<MpnnRun>();
    run->setup(c, A, B, alpha, beta);
}
void LearnedPrecond::apply(double* W, int ldw, int m) { run->apply(net->wbf.as<uint16_t>(), W, ldw, m); }
void LearnedPrecond::apply(float* W, int ldw, int m) { run->apply(net->wbf.as<uint16_t>(), W, ldw, m); }
LearnedPrecond::LearnedPrecond() = default;
LearnedPrecond::~LearnedPrecond() = default;

extern "C" {

int lh_mpnn_param_count(void) { return NPARAM; }

int lh_mpnn_create(lh_ctx* ctx, const float* params, lh_mpnn** out) {
    if (!ctx || !out) return LH_ERR_ARG;
    *out = nullptr;
    LH_API_BEGIN(ctx)
    HIPX(hipSetDevice(ctx->device));
    auto net = std::make_unique<lh_mpnn>();
    net->ctx = ctx;
    net->params.assign(NPARAM, 0.0f);
    if (params) {
        for (int i = 0; i < NPARAM; ++i) net->params[i] = params[i];
    } else {
        // default: small random message weights and a small readout, so that
        // P starts close to Jacobi.  (v = 0 would be exactly Jacobi, but P is
        // quadratic in v and its gradient vanishes there.)
        // The gates start at 1 (Gam = 0) and are trained from there.
        std::mt19937 rng(20260101u);
        std::normal_distribution<float> nd(0.0f, 1.0f);
        for (int i = 0; i < off_gam(0); ++i) net->params[i] = 0.18f * nd(rng);
        for (int f = 0; f < F; ++f) net->params[off_v() + f] = 0.05f * nd(rng);
    }
    mpnn_upload(net.get());
    *out = net.release();
    LH_API_END(ctx)
}

void lh_mpnn_free(lh_mpnn* net) {
    if (!net) return;
    lh_ctx* ctx = net->ctx;  // set at creation
    (void)hipSetDevice(ctx->device);
    if (ctx->mpnn == net) ctx->mpnn = nullptr;
    for (lh_ctx* w : ctx->workers) {
        if (w->mpnn == net) w->mpnn = nullptr;
        (void)hipStreamSynchronize(w->stream);
    }
    (void)hipStreamSynchronize(ctx->stream);
    delete net;
}

int lh_mpnn_set_params(lh_mpnn* net, const float* params) {
    if (!net || !params) return LH_ERR_ARG;
    lh_ctx* ctx = net->ctx;
    LH_API_BEGIN(ctx)
    HIPX(hipSetDevice(ctx->device));
    for (int i = 0; i < NPARAM; ++i) net->params[i] = params[i];
    mpnn_upload(net);
    LH_API_END(ctx)
}

int lh_mpnn_get_params(lh_mpnn* net, float* params) {
    if (!net || !params) return LH_ERR_ARG;
    for (int i = 0; i < NPARAM; ++i) params[i] = net->params[i];
    return LH_OK;
}

int lh_ctx_set_mpnn(lh_ctx* ctx, lh_mpnn* net) {
    if (!ctx) return LH_ERR_ARG;
    if (net && net->ctx != ctx) {
        ctx->last_error = "the network belongs to another context";
        return LH_ERR_ARG;
    }
    ctx->mpnn = net;
    return LH_OK;
}

static void check_mpnn_args(const lh_csr* A, const lh_csr* B, int64_t n) {
    if (B && B->n() != n) LH_THROW(LH_ERR_ARG, "A and B differ in size");
    if (B && !B->is_diag && B->pat != A->pat) LH_THROW(LH_ERR_ARG, "B must be diagonal or share A's pattern");
}

int lh_mpnn_apply(lh_mpnn* net, lh_csr* A, lh_csr* B, double alpha, double beta, const double* R, int nc,
                  double* Z) {
    if (!net || !A || !R || !Z || nc <= 0) return LH_ERR_ARG;
    lh_ctx* ctx = net->ctx;
    LH_API_BEGIN(ctx)
    HIPX(hipSetDevice(ctx->device));
    const int64_t n = A->n();
    check_mpnn_args(A, B, n);
    MpnnRun run;
    run.setup(ctx, A, B, alpha, beta);
    DevBuf dR(n * nc * 8);
    h2d(ctx, dR.p, R, n * nc * 8);
    run.apply(net->wbf.as<uint16_t>(), dR.as<double>(), nc, nc);
    d2h(ctx, Z, dR.p, n * nc * 8);
    LH_API_END(ctx)
}

int lh_mpnn_loss_grad(lh_mpnn* net, lh_csr* A, lh_csr* B, double alpha, double beta, const double* R, int nc,
                      double* loss, double* grad) {
    if (!net || !A || !R || !loss || !grad || nc <= 0) return LH_ERR_ARG;
    lh_ctx* ctx = net->ctx;
    LH_API_BEGIN(ctx)
    HIPX(hipSetDevice(ctx->device));
    const int64_t n = A->n();
    check_mpnn_args(A, B, n);
    MpnnRun run;
    run.setup(ctx, A, B, alpha, beta);
    DevBuf dR(n * nc * 8);
    h2d(ctx, dR.p, R, n * nc * 8);
    mpnn_loss_grad(run, net->wbf.as<uint16_t>(), dR.as<double>(), nc, loss, grad);
    LH_API_END(ctx)
}

int lh_mpnn_time_layer(lh_mpnn* net, lh_csr* A, lh_csr* B, double alpha, double beta, int nc, int iters,
                       int reorder, double* avg_us) {
    if (!net || !A || !avg_us || nc <= 0 || iters <= 0) return LH_ERR_ARG;
    lh_ctx* ctx = net->ctx;
    LH_API_BEGIN(ctx)
    HIPX(hipSetDevice(ctx->device));
    check_mpnn_args(A, B, A->n());
    std::unique_ptr<lh_csr> Ap, Bp;
    if (reorder && (!B || B->is_diag || B->pat == A->pat)) {
        // the eigensolver's locality order, as the solver applies the network
        const auto rop = pattern_reorder(ctx, *A->pat);
        const Reorder& ro = *rop;
        Ap = ro.apply(ctx, A);
        if (B) Bp = ro.apply(ctx, B, B->is_diag ? nullptr : ro.pat);
        A = Ap.get();
        if (B) B = Bp.get();
    }
    MpnnRun run;
    run.setup(ctx, A, B, alpha, beta);
    *avg_us = run.time_layers(net->wbf.as<uint16_t>(), nc, iters);
    LH_API_END(ctx)
}

}  // extern "C"
