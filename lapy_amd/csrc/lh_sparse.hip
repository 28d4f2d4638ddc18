// CSR sparse-matrix x dense-block products (SpMM) on row-major blocks.
//
// Blocks of vectors are stored row-major (n x ld): the ncol values of one
// vertex are contiguous, so the gathered rows X[col, :] of one nonzero are a
// single coalesced load across the lanes that share a matrix row.  A group of
// LPR lanes (16, 32 or 64) owns one matrix row; the row's column indices and
// values are read once per group and broadcast.
//
// Replaces scipy's csc matvec (lapy/solver.py:844,846 and inside eigsh).
#include "lh_internal.hpp"

namespace lh {

template <typename VT, typename XT, int LPR>
__global__ void __launch_bounds__(256)
k_spmm(int64_t n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cols,
       const VT* __restrict__ vals, const XT* __restrict__ X, XT* __restrict__ Y, int ncol, int ldx, int ldy) {
    const int64_t gid = xcd_block() * blockDim.x + threadIdx.x;
    const int64_t row = gid / LPR;
    const int lane = (int)(gid % LPR);
    if (row >= n) return;
    const int32_t r0 = rowptr[row], r1 = rowptr[row + 1];
    for (int c = lane; c < ncol; c += LPR) {
        XT acc = XT(0);
        for (int32_t s = r0; s < r1; ++s) acc += (XT)vals[s] * X[(int64_t)cols[s] * ldx + c];
        Y[row * ldy + c] = acc;
    }
}

// Y1 = A X, Y2 = B X for A, B on one pattern (consistent mass).
template <typename VT, typename XT, int LPR>
__global__ void __launch_bounds__(256)
k_spmm2(int64_t n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cols,
        const VT* __restrict__ va, const VT* __restrict__ vb, const XT* __restrict__ X,
        XT* __restrict__ Y1, XT* __restrict__ Y2, int ncol, int ldx, int ldy) {
    const int64_t gid = xcd_block() * blockDim.x + threadIdx.x;
    const int64_t row = gid / LPR;
    const int lane = (int)(gid % LPR);
    if (row >= n) return;
    const int32_t r0 = rowptr[row], r1 = rowptr[row + 1];
    for (int c = lane; c < ncol; c += LPR) {
        XT a1 = XT(0), a2 = XT(0);
        for (int32_t s = r0; s < r1; ++s) {
            XT x = X[(int64_t)cols[s] * ldx + c];
            a1 += (XT)va[s] * x;
            a2 += (XT)vb[s] * x;
        }
        Y1[row * ldy + c] = a1;
        Y2[row * ldy + c] = a2;
    }
}

// Vectorised forms: each lane owns 16 bytes of the block row (4 floats or 2
// doubles), LPR lanes per matrix row, 64 / LPR rows per wavefront, and two
// gathers in flight per lane.  Used when the widths and offsets allow 16-byte
// aligned accesses.
template <typename VT, typename XT, int LPR>
__global__ void __launch_bounds__(256)
k_spmm_v(int64_t n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cols,
         const VT* __restrict__ vals, const XT* __restrict__ X, XT* __restrict__ Y, int ncol, int ldx, int ldy) {
    using V = typename Vec16<XT>::type;
    const int64_t gid = xcd_block() * blockDim.x + threadIdx.x;
    const int64_t row = gid / LPR;
    const int c0 = (int)(gid % LPR) * Vec16<XT>::n;
    if (row >= n || c0 >= ncol) return;
    const V acc = row_gather<VT, XT>(cols, vals, rowptr[row], rowptr[row + 1], X, ldx, c0);
    *reinterpret_cast<V*>(Y + row * ldy + c0) = acc;
}

template <typename VT, typename XT, int LPR>
__global__ void __launch_bounds__(256)
k_spmm2_v(int64_t n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cols,
          const VT* __restrict__ va, const VT* __restrict__ vb, const XT* __restrict__ X, XT* __restrict__ Y1,
          XT* __restrict__ Y2, int ncol, int ldx, int ldy) {
    using V = typename Vec16<XT>::type;
    const int64_t gid = xcd_block() * blockDim.x + threadIdx.x;
    const int64_t row = gid / LPR;
    const int c0 = (int)(gid % LPR) * Vec16<XT>::n;
    if (row >= n || c0 >= ncol) return;
    V a1 = vzero<XT>(), a2 = vzero<XT>(), b1 = vzero<XT>(), b2 = vzero<XT>();
    const int32_t r1 = rowptr[row + 1];
    int32_t s = rowptr[row];
    for (; s + 1 < r1; s += 2) {
        const int32_t j0 = cols[s], j1 = cols[s + 1];
        const V x0 = *reinterpret_cast<const V*>(X + (int64_t)j0 * ldx + c0);
        const V x1 = *reinterpret_cast<const V*>(X + (int64_t)j1 * ldx + c0);
        vfma<XT>(a1, (XT)va[s], x0);
        vfma<XT>(b1, (XT)vb[s], x0);
        vfma<XT>(a2, (XT)va[s + 1], x1);
        vfma<XT>(b2, (XT)vb[s + 1], x1);
    }
    if (s < r1) {
        const V x0 = *reinterpret_cast<const V*>(X + (int64_t)cols[s] * ldx + c0);
        vfma<XT>(a1, (XT)va[s], x0);
        vfma<XT>(b1, (XT)vb[s], x0);
    }
    vadd<XT>(a1, a2);
    vadd<XT>(b1, b2);
    *reinterpret_cast<V*>(Y1 + row * ldy + c0) = a1;
    *reinterpret_cast<V*>(Y2 + row * ldy + c0) = b1;
}

template <typename XT>
static bool vec_ok(const XT* X, const XT* Y, int ncol, int ldx, int ldy) {
    constexpr int v = Vec16<XT>::n;
    return ncol % v == 0 && ldx % v == 0 && ldy % v == 0 && vec_lanes(ncol, v) >= 4 &&
           (reinterpret_cast<uintptr_t>(X) % 16) == 0 && (reinterpret_cast<uintptr_t>(Y) % 16) == 0;
}

#define LH_VEC_DISPATCH(KERNEL, LANES, ...)                                                          \
    do {                                                                                             \
        const unsigned nbv = nblocks_xcd(n * (LANES), 256);                                          \
        switch (LANES) {                                                                             \
            case 4: KERNEL<VT, XT, 4><<<nbv, 256, 0, ctx->stream>>>(__VA_ARGS__); break;             \
            case 8: KERNEL<VT, XT, 8><<<nbv, 256, 0, ctx->stream>>>(__VA_ARGS__); break;             \
            case 16: KERNEL<VT, XT, 16><<<nbv, 256, 0, ctx->stream>>>(__VA_ARGS__); break;           \
            case 32: KERNEL<VT, XT, 32><<<nbv, 256, 0, ctx->stream>>>(__VA_ARGS__); break;           \
            default: KERNEL<VT, XT, 64><<<nbv, 256, 0, ctx->stream>>>(__VA_ARGS__); break;           \
        }                                                                                            \
    } while (0)

// Y = diag(d) X
template <typename VT, typename XT>
__global__ void k_diag_scale(int64_t n, const VT* __restrict__ d, const XT* __restrict__ X, XT* __restrict__ Y,
                             int ncol, int ldx, int ldy) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t row = gid / ncol;
    const int c = (int)(gid % ncol);
    if (row >= n) return;
    Y[row * ldy + c] = (XT)d[row] * X[row * ldx + c];
}

static int pick_lpr(int ncol) { return ncol > 32 ? 64 : (ncol > 16 ? 32 : 16); }

template <typename VT, typename XT>
void spmm_ld(lh_ctx* ctx, const lh_csr* a, const XT* X, XT* Y, int ncol, int ldx, int ldy) {
    const int64_t n = a->n();
    if (a->is_diag) {
        k_diag_scale<VT, XT><<<nblocks(n * ncol, 256), 256, 0, ctx->stream>>>(n, a->dense_diag.as<VT>(), X, Y, ncol,
                                                                             ldx, ldy);
        LH_CHECK_LAUNCH();
        return;
    }
    const int lpr = pick_lpr(ncol);
    const unsigned nb = nblocks_xcd(n * lpr, 256);
    const int32_t* rp = a->pat->rowptr.as<int32_t>();
    const int32_t* cl = a->pat->cols.as<int32_t>();
    const VT* v = a->vals.as<VT>();
    if (vec_ok<XT>(X, Y, ncol, ldx, ldy) && vec_lanes(ncol, Vec16<XT>::n) <= 64) {
        const int lanes = vec_lanes(ncol, Vec16<XT>::n);
        LH_VEC_DISPATCH(k_spmm_v, lanes, n, rp, cl, v, X, Y, ncol, ldx, ldy);
        LH_CHECK_LAUNCH();
        return;
    }
    if (lpr == 64) k_spmm<VT, XT, 64><<<nb, 256, 0, ctx->stream>>>(n, rp, cl, v, X, Y, ncol, ldx, ldy);
    else if (lpr == 32) k_spmm<VT, XT, 32><<<nb, 256, 0, ctx->stream>>>(n, rp, cl, v, X, Y, ncol, ldx, ldy);
    else k_spmm<VT, XT, 16><<<nb, 256, 0, ctx->stream>>>(n, rp, cl, v, X, Y, ncol, ldx, ldy);
    LH_CHECK_LAUNCH();
}

template <typename VT, typename XT>
void spmm(lh_ctx* ctx, const lh_csr* a, const XT* X, XT* Y, int ncol, int ld) {
    spmm_ld<VT, XT>(ctx, a, X, Y, ncol, ld, ld);
}

template <typename VT, typename XT>
void spmm_pair(lh_ctx* ctx, const lh_csr* a, const lh_csr* b, const XT* X, XT* YA, XT* YB, int ncol, int ldx,
               int ldy) {
    if (!b->is_diag && a->pat == b->pat) {
        const int64_t n = a->n();
        const int lpr = pick_lpr(ncol);
        const unsigned nb = nblocks_xcd(n * lpr, 256);
        const int32_t* rp = a->pat->rowptr.as<int32_t>();
        const int32_t* cl = a->pat->cols.as<int32_t>();
        const VT* va = a->vals.as<VT>();
        const VT* vb = b->vals.as<VT>();
        if (vec_ok<XT>(X, YA, ncol, ldx, ldy) && vec_ok<XT>(X, YB, ncol, ldx, ldy) &&
            vec_lanes(ncol, Vec16<XT>::n) <= 64) {
            const int lanes = vec_lanes(ncol, Vec16<XT>::n);
            LH_VEC_DISPATCH(k_spmm2_v, lanes, n, rp, cl, va, vb, X, YA, YB, ncol, ldx, ldy);
            LH_CHECK_LAUNCH();
            return;
        }
        if (lpr == 64) k_spmm2<VT, XT, 64><<<nb, 256, 0, ctx->stream>>>(n, rp, cl, va, vb, X, YA, YB, ncol, ldx, ldy);
        else if (lpr == 32) k_spmm2<VT, XT, 32><<<nb, 256, 0, ctx->stream>>>(n, rp, cl, va, vb, X, YA, YB, ncol, ldx, ldy);
        else k_spmm2<VT, XT, 16><<<nb, 256, 0, ctx->stream>>>(n, rp, cl, va, vb, X, YA, YB, ncol, ldx, ldy);
        LH_CHECK_LAUNCH();
        return;
    }
    spmm_ld<VT, XT>(ctx, a, X, YA, ncol, ldx, ldy);
    spmm_ld<VT, XT>(ctx, b, X, YB, ncol, ldx, ldy);
}

template void spmm<double, double>(lh_ctx*, const lh_csr*, const double*, double*, int, int);
template void spmm<float, double>(lh_ctx*, const lh_csr*, const double*, double*, int, int);
template void spmm<float, float>(lh_ctx*, const lh_csr*, const float*, float*, int, int);
template void spmm<double, float>(lh_ctx*, const lh_csr*, const float*, float*, int, int);
template void spmm_ld<double, double>(lh_ctx*, const lh_csr*, const double*, double*, int, int, int);
template void spmm_ld<float, float>(lh_ctx*, const lh_csr*, const float*, float*, int, int, int);
template void spmm_pair<double, double>(lh_ctx*, const lh_csr*, const lh_csr*, const double*, double*, double*, int,
                                        int, int);
template void spmm_pair<float, float>(lh_ctx*, const lh_csr*, const lh_csr*, const float*, float*, float*, int, int,
                                      int);

__global__ void k_f64_to_f32(const double* __restrict__ s, float* __restrict__ d, int64_t cnt) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) d[i] = (float)s[i];
}
void copy_block_f64_f32(lh_ctx* ctx, const double* s, float* d, int64_t cnt) {
    if (cnt <= 0) return;
    k_f64_to_f32<<<nblocks(cnt, 256), 256, 0, ctx->stream>>>(s, d, cnt);
    LH_CHECK_LAUNCH();
}

std::unique_ptr<lh_csr> csr_to_f32(lh_ctx* ctx, const lh_csr* s) {
    if (s->dtype != LH_F64) LH_THROW(LH_ERR_ARG, "expects an FP64 matrix");
    auto d = std::make_unique<lh_csr>();
    d->ctx = ctx;
    d->pat = s->pat;
    d->dtype = LH_F32;
    d->is_diag = s->is_diag;
    d->vals.alloc(s->nnz() * 4);
    copy_block_f64_f32(ctx, s->vals.as<double>(), d->vals.as<float>(), s->nnz());
    if (s->is_diag) refresh_dense_diag(ctx, d.get());
    return d;
}

template <typename T>
__global__ void k_fill_pattern(int64_t cnt, T* __restrict__ x) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) x[i] = T(1) + T(i % 7) * T(0.125);
}

}  // namespace lh

using namespace lh;

extern "C" int lh_time_spmm(lh_ctx* ctx, lh_csr* A, lh_csr* B, int ncol, int precision, int iters, double* avg_us) {
    if (!ctx || !A || !B || !avg_us || iters <= 0 || ncol <= 0) return LH_ERR_ARG;
    LH_API_BEGIN(ctx)
    HIPX(hipSetDevice(ctx->device));
    const int64_t n = A->n();
    const size_t es = precision == LH_PREC_FP64 ? 8 : 4;
    DevBuf X(n * ncol * es), YA(n * ncol * es), YB(n * ncol * es);
    // FP32 SpMM needs FP32 values: convert on the fly for the timing copy
    std::unique_ptr<lh_csr> A32, B32;
    const lh_csr *pa = A, *pb = B;
    if (precision != LH_PREC_FP64) {
        A32 = csr_to_f32(ctx, A);
        B32 = csr_to_f32(ctx, B);
        pa = A32.get();
        pb = B32.get();
    }
    auto launch = [&]() {
        if (precision == LH_PREC_FP64) spmm_pair<double, double>(ctx, pa, pb, X.as<double>(), YA.as<double>(), YB.as<double>(), ncol, ncol, ncol);
        else spmm_pair<float, float>(ctx, pa, pb, X.as<float>(), YA.as<float>(), YB.as<float>(), ncol, ncol, ncol);
    };
    if (precision == LH_PREC_FP64) k_fill_pattern<double><<<nblocks(n * ncol, 256), 256, 0, ctx->stream>>>(n * ncol, X.as<double>());
    else k_fill_pattern<float><<<nblocks(n * ncol, 256), 256, 0, ctx->stream>>>(n * ncol, X.as<float>());
    launch();
    hipEvent_t e0, e1;
    HIPX(hipEventCreate(&e0));
    HIPX(hipEventCreate(&e1));
    HIPX(hipEventRecord(e0, ctx->stream));
    for (int i = 0; i < iters; ++i) launch();
    HIPX(hipEventRecord(e1, ctx->stream));
    HIPX(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPX(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_us = 1000.0 * ms / iters;
    LH_API_END(ctx)
}
