// Learned message-passing preconditioner (no counterpart in the reference,
// which factorises A - sigma B with SuperLU, lapy/solver.py:707-715).
//
// For the SPD operator K = beta A + alpha B and residual columns R (n x c,
// row-major, leading dimension ld):
//
//     s = 1 / sqrt(diag K),  W^ = S (K - diag K) S   (FP32, on K's pattern)
//     Q x:   H0 = x u^T, H_{l+1} = H_l Ws_l^T + (W^ H_l) Wn_l^T (l = 0..2,
//            F = 32), Q x = H3 v
//     Q^T y: G3 = y v^T, G_l = G_{l+1} Ws_l + W^ (G_{l+1} Wn_l), Q^T y = G0 u
//     z = s (x + Q^T Q x), x = s r        P = S (I + Q^T Q) S: SPD, v = 0 is
//                                          exactly Jacobi
//
// Weights in bfloat16, arithmetic in FP32, hidden states FP32 laid out
// [vertex][column][feature] (one 128-byte row per vertex and column).  A
// group of 32 lanes owns one (vertex, column) pair: lane f gathers feature f
// of the neighbours (coalesced 128-byte rows), stages h and the aggregate in
// LDS, and forms output feature f from the 32 x 32 weights staged in LDS.
//
// Q^T runs on the backward kernels (the adjoint of a linear layer is its
// input-gradient pass).  Training: L = (1/c) sum_c ||S (K z_c - r_c)||^2 /
// ||S r_c||^2; with a = Q x and g_b = S dL/dz the gradient is the sum of two
// back-propagations through Q, (x, Q g_b) and (g_b, a), because
// <g_b, Q^T a> = <Q g_b, a>.  Weight gradients are accumulated per lane
// over the pairs a group visits, reduced over the groups of a workgroup in a
// fixed order and over the workgroups by a separate kernel (FP64), so the
// gradient is deterministic.
#include "lh_dense.hpp"

#include <cmath>
#include <cstring>
#include <random>

struct lh_mpnn {
    lh_ctx* ctx = nullptr;
    std::vector<float> params;  // host master copy
    DevBuf wbf;                 // device bfloat16 copy
};

namespace lh {

namespace {

constexpr int F = 32;
constexpr int NL = 3;
constexpr int NPARAM = 2 * F + 2 * NL * F * F;
constexpr int GROUPS = 8;  // (vertex, column) pairs per 256-thread workgroup

// parameter offsets in the flat vector: u, (Ws_l, Wn_l) for l = 0..2, v
__host__ __device__ constexpr int off_u() { return 0; }
__host__ __device__ constexpr int off_ws(int l) { return F + l * 2 * F * F; }
__host__ __device__ constexpr int off_wn(int l) { return F + l * 2 * F * F + F * F; }
__host__ __device__ constexpr int off_v() { return F + NL * 2 * F * F; }

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

uint16_t f2bf(float f) {  // round to nearest even
    uint32_t u;
    std::memcpy(&u, &f, 4);
    const uint32_t lsb = (u >> 16) & 1u;
    return (uint16_t)((u + 0x7FFFu + lsb) >> 16);
}

// s_i = 1/sqrt(K_ii) (0 for K_ii <= 0)
__global__ void k_mp_scale(int64_t n, const int32_t* __restrict__ diagpos, const double* __restrict__ kv,
                           float* __restrict__ s) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double d = diagpos[i] >= 0 ? kv[diagpos[i]] : 0.0;
    s[i] = d > 0.0 ? (float)(1.0 / sqrt(d)) : 0.0f;
}
// w_slot = s_i K_ij s_j off the diagonal, 0 on it
__global__ void k_mp_weights(int64_t n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cols,
                             const double* __restrict__ kv, const float* __restrict__ s, float* __restrict__ w) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double si = s[i];
    for (int32_t p = rowptr[i]; p < rowptr[i + 1]; ++p) {
        const int32_t j = cols[p];
        w[p] = (j == (int32_t)i) ? 0.0f : (float)(si * kv[p] * (double)s[j]);
    }
}

// H0[pc][f] = x_pc u_f with x = X (leading dimension ldx), scaled by s when
// s is given
__global__ void k_mp_lift(int64_t n, int nc, const double* __restrict__ X, int ldx, const float* __restrict__ s,
                          const uint16_t* __restrict__ wb, float* __restrict__ H0) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n * nc * F) return;
    const int f = (int)(gid % F);
    const int64_t pc = gid / F;
    const int64_t i = pc / nc;
    const int c = (int)(pc - i * nc);
    const float x = (float)X[i * ldx + c] * (s ? s[i] : 1.0f);
    H0[gid] = x * bf2f(wb[off_u() + f]);
}

__device__ __forceinline__ void load_mats(float* sWs, float* sWn, const uint16_t* __restrict__ wb, int l) {
    for (int t = threadIdx.x; t < F * F; t += blockDim.x) {
        sWs[t] = bf2f(wb[off_ws(l) + t]);
        sWn[t] = bf2f(wb[off_wn(l) + t]);
    }
}

// one layer: Hout = Hin Ws^T + (W^ Hin) Wn^T; optional G = W^ Hin
__global__ void __launch_bounds__(256)
k_mp_layer(int64_t n, int nc, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cols,
           const float* __restrict__ w, const float* __restrict__ Hin, const uint16_t* __restrict__ wb, int l,
           float* __restrict__ Hout, float* __restrict__ Gout) {
    __shared__ float sWs[F * F], sWn[F * F];
    __shared__ float sh[GROUPS][F], sg[GROUPS][F];
    load_mats(sWs, sWn, wb, l);
    const int grp = threadIdx.x / F, f = threadIdx.x % F;
    const int64_t npair = n * nc;
    for (int64_t base = (int64_t)blockIdx.x * GROUPS; base < npair; base += (int64_t)gridDim.x * GROUPS) {
        const int64_t pc = base + grp;
        const bool ok = pc < npair;
        float h = 0.f, g = 0.f;
        if (ok) {
            const int64_t i = pc / nc;
            const int c = (int)(pc - i * nc);
            h = Hin[pc * F + f];
            for (int32_t p = rowptr[i]; p < rowptr[i + 1]; ++p) {
                const float wij = w[p];
                if (wij != 0.f) g += wij * Hin[((int64_t)cols[p] * nc + c) * F + f];
            }
        }
        __syncthreads();  // the previous pair's staging has been read
        sh[grp][f] = h;
        sg[grp][f] = g;
        __syncthreads();
        if (ok) {
            float o = 0.f;
#pragma unroll 8
            for (int k = 0; k < F; ++k) o += sWs[f * F + k] * sh[grp][k] + sWn[f * F + k] * sg[grp][k];
            Hout[pc * F + f] = o;
            if (Gout) Gout[pc * F + f] = g;
        }
    }
}

// Y_pc = H[pc] . wvec with wvec = u (woff = 0) or v (woff = off_v())
__global__ void __launch_bounds__(256)
k_mp_readout(int64_t n, int nc, const float* __restrict__ H, const uint16_t* __restrict__ wb, int woff,
             double* __restrict__ Y) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t pc = gid / F;
    const int f = (int)(gid % F);
    const bool ok = pc < n * nc;
    float y = ok ? H[pc * F + f] * bf2f(wb[woff + f]) : 0.f;
    for (int off = 16; off > 0; off >>= 1) y += __shfl_xor(y, off, 64);
    if (ok && f == 0) Y[pc] = (double)y;
}

// z = s (s r + b), written over Z (may alias R: same element, same thread)
__global__ void k_mp_combine(int64_t n, int nc, const double* R, int ldr, const float* __restrict__ s,
                             const double* __restrict__ Bv, double* Z, int ldz) {
    const int64_t pc = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pc >= n * nc) return;
    const int64_t i = pc / nc;
    const int c = (int)(pc - i * nc);
    const float si = s[i];
    const float x = si * (float)R[i * ldr + c];
    Z[i * ldz + c] = (double)(si * (x + (float)Bv[pc]));
}

// dH3[pc][f] = dy_pc v_f; part (nullable): per-workgroup dv = sum dy H3
__global__ void __launch_bounds__(256)
k_mp_out_bwd(int64_t n, int nc, const double* __restrict__ dY, const float* __restrict__ H3,
             const uint16_t* __restrict__ wb, float* __restrict__ dH3, double* __restrict__ part) {
    __shared__ float red[GROUPS][F];
    const int grp = threadIdx.x / F, f = threadIdx.x % F;
    const float vf = bf2f(wb[off_v() + f]);
    const int64_t npair = n * nc;
    float acc = 0.f;
    for (int64_t pc = (int64_t)blockIdx.x * GROUPS + grp; pc < npair; pc += (int64_t)gridDim.x * GROUPS) {
        const float dy = (float)dY[pc];
        dH3[pc * F + f] = dy * vf;
        if (part) acc += dy * H3[pc * F + f];
    }
    if (!part) return;
    red[grp][f] = acc;
    __syncthreads();
    if (grp == 0) {
        double t = 0.0;
        for (int g = 0; g < GROUPS; ++g) t += (double)red[g][f];
        part[(int64_t)blockIdx.x * F + f] = t;
    }
}

// local part of the layer backward: dHloc = dHout Ws, T = dHout Wn, and (when
// part is given) the weight-gradient partials dWs += dHout^T Hin, dWn +=
// dHout^T G
__global__ void __launch_bounds__(256)
k_mp_layer_bwd(int64_t n, int nc, const float* __restrict__ dHout, const float* __restrict__ Hin,
               const float* __restrict__ G, const uint16_t* __restrict__ wb, int l, float* __restrict__ dHloc,
               float* __restrict__ T, double* __restrict__ part) {
    __shared__ float sWs[F * F], sWn[F * F];
    __shared__ float sd[GROUPS][F], shh[GROUPS][F], sgg[GROUPS][F];
    __shared__ float acc_s[F * F], acc_n[F * F];
    load_mats(sWs, sWn, wb, l);
    const bool grads = part != nullptr;
    for (int t = threadIdx.x; t < F * F; t += blockDim.x) { acc_s[t] = 0.f; acc_n[t] = 0.f; }
    const int grp = threadIdx.x / F, f = threadIdx.x % F;
    const int64_t npair = n * nc;
    float as[F], an[F];
#pragma unroll
    for (int k = 0; k < F; ++k) { as[k] = 0.f; an[k] = 0.f; }
    for (int64_t base = (int64_t)blockIdx.x * GROUPS; base < npair; base += (int64_t)gridDim.x * GROUPS) {
        const int64_t pc = base + grp;
        const bool ok = pc < npair;
        __syncthreads();
        sd[grp][f] = ok ? dHout[pc * F + f] : 0.f;
        if (grads) {
            shh[grp][f] = ok ? Hin[pc * F + f] : 0.f;
            sgg[grp][f] = ok ? G[pc * F + f] : 0.f;
        }
        __syncthreads();
        if (ok) {
            float a = 0.f, b = 0.f;
#pragma unroll 8
            for (int q = 0; q < F; ++q) {
                a += sd[grp][q] * sWs[q * F + f];
                b += sd[grp][q] * sWn[q * F + f];
            }
            dHloc[pc * F + f] = a;
            T[pc * F + f] = b;
            if (grads) {
                const float df = sd[grp][f];
#pragma unroll
                for (int k = 0; k < F; ++k) {
                    as[k] += df * shh[grp][k];
                    an[k] += df * sgg[grp][k];
                }
            }
        }
    }
    if (!grads) return;
    // groups add their rows into the workgroup accumulator one after another
    for (int g = 0; g < GROUPS; ++g) {
        __syncthreads();
        if (grp == g)
#pragma unroll
            for (int k = 0; k < F; ++k) {
                acc_s[f * F + k] += as[k];
                acc_n[f * F + k] += an[k];
            }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < F * F; t += blockDim.x) {
        part[(int64_t)blockIdx.x * 2 * F * F + t] = acc_s[t];
        part[(int64_t)blockIdx.x * 2 * F * F + F * F + t] = acc_n[t];
    }
}

// dHin = dHloc + W^ T (W^ symmetric)
__global__ void __launch_bounds__(256)
k_mp_agg_bwd(int64_t n, int nc, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cols,
             const float* __restrict__ w, const float* __restrict__ dHloc, const float* __restrict__ T,
             float* __restrict__ dHin) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n * nc * F) return;
    const int f = (int)(gid % F);
    const int64_t pc = gid / F;
    const int64_t i = pc / nc;
    const int c = (int)(pc - i * nc);
    float g = 0.f;
    for (int32_t p = rowptr[i]; p < rowptr[i + 1]; ++p) {
        const float wij = w[p];
        if (wij != 0.f) g += wij * T[((int64_t)cols[p] * nc + c) * F + f];
    }
    dHin[gid] = dHloc[gid] + g;
}

// du partial: sum x_pc dH0[pc][f], x = X (ldx) scaled by s when given
__global__ void __launch_bounds__(256)
k_mp_lift_bwd(int64_t n, int nc, const double* __restrict__ X, int ldx, const float* __restrict__ s,
              const float* __restrict__ dH0, double* __restrict__ part) {
    __shared__ float red[GROUPS][F];
    const int grp = threadIdx.x / F, f = threadIdx.x % F;
    const int64_t npair = n * nc;
    float acc = 0.f;
    for (int64_t pc = (int64_t)blockIdx.x * GROUPS + grp; pc < npair; pc += (int64_t)gridDim.x * GROUPS) {
        const int64_t i = pc / nc;
        const int c = (int)(pc - i * nc);
        const float x = (float)X[i * ldx + c] * (s ? s[i] : 1.0f);
        acc += x * dH0[pc * F + f];
    }
    red[grp][f] = acc;
    __syncthreads();
    if (grp == 0) {
        double t = 0.0;
        for (int g = 0; g < GROUPS; ++g) t += (double)red[g][f];
        part[(int64_t)blockIdx.x * F + f] = t;
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
__global__ void k_mp_reduce(int nb, int len, const double* __restrict__ part, double* __restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= len) return;
    double t = 0.0;
    for (int b = 0; b < nb; ++b) t += part[(int64_t)b * len + j];
    out[j] = t;
}

constexpr int kPairBlocks = 1024;  // workgroups of the grid-stride pair loops

}  // namespace

// The operator-dependent state (K, s, W^) and the work buffers of the
// network; reused across applications with the same K.
struct MpnnRun {
    lh_ctx* ctx = nullptr;
    std::shared_ptr<Pattern> pat;
    int64_t n = 0;
    DevBuf kv, s, w;
    DevBuf HA[NL + 1], GA[NL], HB[NL + 1], GB[NL];  // activations of two forward passes
    DevBuf dH, dHloc, T;                            // adjoint / back-propagation
    DevBuf fa, fb;                                  // scalar fields (n x nc, FP64)

    void setup(lh_ctx* c, const lh_csr* A, const lh_csr* B, double alpha, double beta) {
        ctx = c;
        pat = A->pat;
        n = A->n();
        form_shifted(ctx, A, B, alpha, beta, kv);
        s.ensure(n * 4);
        w.ensure(pat->nnz * 4);
        k_mp_scale<<<nblocks(n, 256), 256, 0, ctx->stream>>>(n, pat->diagpos.as<int32_t>(), kv.as<double>(),
                                                            s.as<float>());
        k_mp_weights<<<nblocks(n, 256), 256, 0, ctx->stream>>>(n, pat->rowptr.as<int32_t>(), pat->cols.as<int32_t>(),
                                                              kv.as<double>(), s.as<float>(), w.as<float>());
        LH_CHECK_LAUNCH();
    }
    size_t hbytes(int nc) const { return (size_t)n * nc * F * 4; }
    // Y (dense n x nc) = Q x, x = X (ldx), times s when scaled.  store: keep
    // every layer's input and aggregate in H, G; else ping-pong H[0], H[1]
    void q_forward(const uint16_t* wb, const double* X, int ldx, bool scaled, int nc, DevBuf* H, DevBuf* G, bool store,
                   double* Y) {
        for (int l = 0; l <= (store ? NL : 1); ++l) H[l].ensure(hbytes(nc));
        if (store)
            for (int l = 0; l < NL; ++l) G[l].ensure(hbytes(nc));
        const int64_t tot = n * nc * F;
        k_mp_lift<<<nblocks(tot, 256), 256, 0, ctx->stream>>>(n, nc, X, ldx, scaled ? s.as<float>() : nullptr, wb,
                                                              H[0].as<float>());
        for (int l = 0; l < NL; ++l) {
            const float* hin = H[store ? l : (l & 1)].as<float>();
            float* hout = H[store ? l + 1 : ((l + 1) & 1)].as<float>();
            k_mp_layer<<<kPairBlocks, 256, 0, ctx->stream>>>(n, nc, pat->rowptr.as<int32_t>(),
                                                           pat->cols.as<int32_t>(), w.as<float>(), hin, wb, l, hout,
                                                           store ? G[l].as<float>() : nullptr);
        }
        k_mp_readout<<<nblocks(tot, 256), 256, 0, ctx->stream>>>(n, nc, H[store ? NL : (NL & 1)].as<float>(), wb,
                                                                off_v(), Y);
        LH_CHECK_LAUNCH();
    }
    // out (dense n x nc) = Q^T Y
    void q_adjoint(const uint16_t* wb, const double* Y, int nc, double* out) {
        dH.ensure(hbytes(nc));
        dHloc.ensure(hbytes(nc));
        T.ensure(hbytes(nc));
        const int64_t tot = n * nc * F;
        k_mp_out_bwd<<<kPairBlocks, 256, 0, ctx->stream>>>(n, nc, Y, nullptr, wb, dH.as<float>(), nullptr);
        for (int l = NL - 1; l >= 0; --l) {
            k_mp_layer_bwd<<<kPairBlocks, 256, 0, ctx->stream>>>(n, nc, dH.as<float>(), nullptr, nullptr, wb, l,
                                                               dHloc.as<float>(), T.as<float>(), nullptr);
            k_mp_agg_bwd<<<nblocks(tot, 256), 256, 0, ctx->stream>>>(n, nc, pat->rowptr.as<int32_t>(),
                                                                    pat->cols.as<int32_t>(), w.as<float>(),
                                                                    dHloc.as<float>(), T.as<float>(), dH.as<float>());
        }
        k_mp_readout<<<nblocks(tot, 256), 256, 0, ctx->stream>>>(n, nc, dH.as<float>(), wb, off_u(), out);
        LH_CHECK_LAUNCH();
    }
    // gdev (NPARAM doubles) = gradient of <dY, Q x> over the parameters, with
    // the stored activations H, G of input x = X (ldx) (times s when scaled)
    void q_backprop(const uint16_t* wb, const double* X, int ldx, bool scaled, int nc, DevBuf* H, DevBuf* G,
                    const double* dY, DevBuf& partv, DevBuf& partw, double* gdev) {
        dH.ensure(hbytes(nc));
        dHloc.ensure(hbytes(nc));
        T.ensure(hbytes(nc));
        const int64_t tot = n * nc * F;
        k_mp_out_bwd<<<kPairBlocks, 256, 0, ctx->stream>>>(n, nc, dY, H[NL].as<float>(), wb, dH.as<float>(),
                                                           partv.as<double>());
        k_mp_reduce<<<1, 64, 0, ctx->stream>>>(kPairBlocks, F, partv.as<double>(), gdev + off_v());
        for (int l = NL - 1; l >= 0; --l) {
            k_mp_layer_bwd<<<kPairBlocks, 256, 0, ctx->stream>>>(n, nc, dH.as<float>(), H[l].as<float>(),
                                                               G[l].as<float>(), wb, l, dHloc.as<float>(),
                                                               T.as<float>(), partw.as<double>());
            k_mp_reduce<<<nblocks(2 * F * F, 256), 256, 0, ctx->stream>>>(kPairBlocks, 2 * F * F, partw.as<double>(),
                                                                          gdev + off_ws(l));
            k_mp_agg_bwd<<<nblocks(tot, 256), 256, 0, ctx->stream>>>(n, nc, pat->rowptr.as<int32_t>(),
                                                                    pat->cols.as<int32_t>(), w.as<float>(),
                                                                    dHloc.as<float>(), T.as<float>(), dH.as<float>());
        }
        k_mp_lift_bwd<<<kPairBlocks, 256, 0, ctx->stream>>>(n, nc, X, ldx, scaled ? s.as<float>() : nullptr,
                                                            dH.as<float>(), partv.as<double>());
        k_mp_reduce<<<1, 64, 0, ctx->stream>>>(kPairBlocks, F, partv.as<double>(), gdev + off_u());
        LH_CHECK_LAUNCH();
    }
    // W (n x m, ldw) <- P W = S (I + Q^T Q) S W, in place
    void apply(const uint16_t* wb, double* W, int ldw, int m) {
        fa.ensure((size_t)n * m * 8);
        fb.ensure((size_t)n * m * 8);
        q_forward(wb, W, ldw, true, m, HA, GA, false, fa.as<double>());
        q_adjoint(wb, fa.as<double>(), m, fb.as<double>());
        k_mp_combine<<<nblocks(n * m, 256), 256, 0, ctx->stream>>>(n, m, W, ldw, s.as<float>(), fb.as<double>(), W,
                                                                  ldw);
        LH_CHECK_LAUNCH();
    }
};

// loss and gradient (host) for residual columns R (device, n x nc, dense)
void mpnn_loss_grad(MpnnRun& run, const uint16_t* wb, const double* R, int nc, double* loss, double* grad) {
    lh_ctx* ctx = run.ctx;
    const int64_t n = run.n;
    DevBuf a(n * nc * 8), b(n * nc * 8), Z(n * nc * 8), KZ(n * nc * 8), Tt(n * nc * 8), SR(n * nc * 8),
        gb(n * nc * 8), ga(n * nc * 8);
    // z = s (x + Q^T Q x), x = s r
    run.q_forward(wb, R, nc, true, nc, run.HA, run.GA, true, a.as<double>());
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
    // g_a = Q g_b (activations kept for the second back-propagation)
    run.q_forward(wb, gb.as<double>(), nc, false, nc, run.HB, run.GB, true, ga.as<double>());
    // dL/dtheta = grad <g_a, Q x> + grad <a, Q g_b>
    DevBuf partv((size_t)kPairBlocks * F * 8), partw((size_t)kPairBlocks * 2 * F * F * 8), g1(NPARAM * 8),
        g2(NPARAM * 8);
    run.q_backprop(wb, R, nc, true, nc, run.HA, run.GA, ga.as<double>(), partv, partw, g1.as<double>());
    run.q_backprop(wb, gb.as<double>(), nc, false, nc, run.HB, run.GB, a.as<double>(), partv, partw, g2.as<double>());
    std::vector<double> h1(NPARAM), h2(NPARAM);
    d2h(ctx, h1.data(), g1.p, NPARAM * 8);
    d2h(ctx, h2.data(), g2.p, NPARAM * 8);
    for (int i = 0; i < NPARAM; ++i) grad[i] = h1[i] + h2[i];
}

int mpnn_param_count() { return NPARAM; }

void mpnn_upload(lh_mpnn* net) {
    std::vector<uint16_t> b(NPARAM);
    for (int i = 0; i < NPARAM; ++i) b[i] = f2bf(net->params[i]);
    net->wbf.persistent = true;
    net->wbf.ensure(NPARAM * 2);
    h2d(net->ctx, net->wbf.p, b.data(), NPARAM * 2);
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
        std::mt19937 rng(20260101u);
        std::normal_distribution<float> nd(0.0f, 1.0f);
        for (int i = 0; i < NPARAM; ++i) net->params[i] = 0.18f * nd(rng);
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

}  // extern "C"
