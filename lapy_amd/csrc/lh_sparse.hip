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
    if (ncol <= 2) {
        // one or two vectors (SpMV, power iteration): one thread per row
        k_spmm<VT, XT, 1><<<nblocks_xcd(n, 256), 256, 0, ctx->stream>>>(n, rp, cl, v, X, Y, ncol, ldx, ldy);
    } else if (lpr == 64) {
        k_spmm<VT, XT, 64><<<nb, 256, 0, ctx->stream>>>(n, rp, cl, v, X, Y, ncol, ldx, ldy);
    } else if (lpr == 32) {
        k_spmm<VT, XT, 32><<<nb, 256, 0, ctx->stream>>>(n, rp, cl, v, X, Y, ncol, ldx, ldy);
    } else {
        k_spmm<VT, XT, 16><<<nb, 256, 0, ctx->stream>>>(n, rp, cl, v, X, Y, ncol, ldx, ldy);
    }
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

// ---- ELL forms ------------------------------------------------------------
// non-temporal 16-byte store of a block-row vector
__device__ __forceinline__ void nt_store(const double2& v, double* p) {
    typedef double d2n __attribute__((ext_vector_type(2)));
    __builtin_nontemporal_store((d2n){v.x, v.y}, reinterpret_cast<d2n*>(p));
}
__device__ __forceinline__ void nt_store(const float4& v, float* p) {
    typedef float f4n __attribute__((ext_vector_type(4)));
    __builtin_nontemporal_store((f4n){v.x, v.y, v.z, v.w}, reinterpret_cast<f4n*>(p));
}

// TWO: YA = A X and YB = B X in one pass; else Y = M X with M's values in va
template <typename VT, typename XT, int LPR, bool TWO>
__global__ void __launch_bounds__(256)
k_spmm_e(int64_t n, int W, const int32_t* __restrict__ ecols, const VT* __restrict__ va, const VT* __restrict__ vb,
         const XT* __restrict__ X, XT* __restrict__ YA, XT* __restrict__ YB, int ncol, int ldx, int ldy) {
    using V = typename Vec16<XT>::type;
    const int64_t gid = xcd_block() * blockDim.x + threadIdx.x;
    int64_t row = gid / LPR;
    // one row per wavefront: the row, hence its slots, is wave-uniform and
    // the slots come through the scalar cache (loads only)
    if (LPR == 64) row = __builtin_amdgcn_readfirstlane((int)row);
    const int c0 = (int)(gid % LPR) * Vec16<XT>::n;
    if (row >= n || c0 >= ncol) return;
    const int32_t* cr = ecols + row * W;
    const VT* ar = va + row * W;
    const VT* br = TWO ? vb + row * W : nullptr;
    V a1 = vzero<XT>(), a2 = vzero<XT>(), b1 = vzero<XT>(), b2 = vzero<XT>();
    for (int s = 0; s < W; s += 8) {
        const int4 ja = *reinterpret_cast<const int4*>(cr + s);
        const int4 jb = *reinterpret_cast<const int4*>(cr + s + 4);
        const int32_t j[8] = {ja.x, ja.y, ja.z, ja.w, jb.x, jb.y, jb.z, jb.w};
        V x[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) x[q] = *reinterpret_cast<const V*>(X + (int64_t)j[q] * ldx + c0);
#pragma unroll
        for (int q = 0; q < 8; q += 2) {
            vfma<XT>(a1, (XT)ar[s + q], x[q]);
            vfma<XT>(a2, (XT)ar[s + q + 1], x[q + 1]);
            if (TWO) {
                vfma<XT>(b1, (XT)br[s + q], x[q]);
                vfma<XT>(b2, (XT)br[s + q + 1], x[q + 1]);
            }
        }
    }
    // the products are not read again by this pass: non-temporal stores
    vadd<XT>(a1, a2);
    nt_store(a1, YA + row * ldy + c0);
    if (TWO) {
        vadd<XT>(b1, b2);
        nt_store(b1, YB + row * ldy + c0);
    }
}

// longest row of a pattern (into *w, which starts at 0)
__global__ void k_rowlen_max(int64_t n, const int32_t* __restrict__ rowptr, int* __restrict__ w) {
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

template <typename VT>
__global__ void k_csr_to_ell(int64_t n, int W, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cols,
                             const VT* __restrict__ a, const VT* __restrict__ b, int32_t* __restrict__ ecols,
                             VT* __restrict__ ea, VT* __restrict__ eb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t r0 = rowptr[i], len = rowptr[i + 1] - r0;
    for (int q = 0; q < W; ++q) {
        const bool in = q < len;
        ecols[i * W + q] = in ? cols[r0 + q] : (int32_t)i;
        ea[i * W + q] = in ? a[r0 + q] : VT(0);
        if (b) eb[i * W + q] = in ? b[r0 + q] : VT(0);
    }
}

void Ell::build(lh_ctx* ctx, const lh_csr* a, const lh_csr* b) {
    W = 0;
    two = false;
    if (a->is_diag) return;
    if (b && (b->is_diag || b->pat != a->pat || b->dtype != a->dtype)) b = nullptr;
    n = a->n();
    nnz = a->nnz();
    dtype = a->dtype;
    const Pattern& p = *a->pat;
    DevBuf wmax(8);
    memset_d(ctx, wmax.p, 0, 4);
    k_rowlen_max<<<256, 256, 0, ctx->stream>>>(n, p.rowptr.as<int32_t>(), wmax.as<int>());
    LH_CHECK_LAUNCH();
    int w = 0;
    d2h(ctx, &w, wmax.p, 4);
    w = (w + 7) / 8 * 8;
    if (w <= 0 || w > kMaxW) return;
    const size_t es = dtype == LH_F64 ? 8 : 4;
    cols.alloc((size_t)n * w * 4);
    va.alloc((size_t)n * w * es);
    if (b) vb.alloc((size_t)n * w * es);
    if (dtype == LH_F64)
        k_csr_to_ell<double><<<nblocks(n, 256), 256, 0, ctx->stream>>>(
            n, w, p.rowptr.as<int32_t>(), p.cols.as<int32_t>(), a->vals.as<double>(),
            b ? b->vals.as<double>() : nullptr, cols.as<int32_t>(), va.as<double>(), b ? vb.as<double>() : nullptr);
    else
        k_csr_to_ell<float><<<nblocks(n, 256), 256, 0, ctx->stream>>>(
            n, w, p.rowptr.as<int32_t>(), p.cols.as<int32_t>(), a->vals.as<float>(),
            b ? b->vals.as<float>() : nullptr, cols.as<int32_t>(), va.as<float>(), b ? vb.as<float>() : nullptr);
    LH_CHECK_LAUNCH();
    W = w;
    two = b != nullptr;
}

int tuned_spmm_block(const lh_ctx* ctx, int64_t n, int64_t nnz) {
    if (!ctx->tune) return 0;
    std::lock_guard<std::mutex> lk(ctx->tune->mu);
    for (const SpmmTune& t : ctx->tune->spmm)
        if (t.n == n && t.nnz == nnz) return t.block;
    return 0;
}

template <typename VT, typename XT>
bool spmm_ell(lh_ctx* ctx, const Ell& e, const XT* X, XT* YA, XT* YB, int ncol, int ldx, int ldy, int which,
              int block) {
    if (e.W == 0 || (which == 1 && !e.two)) return false;
    const bool pair = which == 0 && e.two && YB != nullptr;
    XT* y = (which == 1) ? YB : YA;
    auto al = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) % 16) == 0; };
    constexpr int v = Vec16<XT>::n;
    const bool vec = ncol % v == 0 && ldx % v == 0 && ldy % v == 0 && vec_lanes(ncol, v) <= 64 && al(X) && al(y) &&
                     (!pair || al(YB));
    if (!vec) return false;
    const int64_t n = e.n;
    const int32_t* ec = e.cols.as<int32_t>();
    const VT* a = (which == 1) ? e.vb.as<VT>() : e.va.as<VT>();
    const VT* b = e.vb.as<VT>();
    const int lanes = vec_lanes(ncol, v);
    if (block == 0) block = tuned_spmm_block(ctx, n, e.nnz);
    if (block != 64 && block != 128) block = 256;
    const unsigned nb = nblocks_xcd(n * lanes, block);
#define LH_ELL_LAUNCH(L)                                                                                     \
    do {                                                                                                     \
        if (pair) k_spmm_e<VT, XT, L, true><<<nb, block, 0, ctx->stream>>>(n, e.W, ec, a, b, X, YA, YB, ncol, ldx, ldy); \
        else k_spmm_e<VT, XT, L, false><<<nb, block, 0, ctx->stream>>>(n, e.W, ec, a, nullptr, X, y, nullptr, ncol, ldx, ldy); \
    } while (0)
    switch (lanes) {
        case 1: LH_ELL_LAUNCH(1); break;
        case 2: LH_ELL_LAUNCH(2); break;
        case 4: LH_ELL_LAUNCH(4); break;
        case 8: LH_ELL_LAUNCH(8); break;
        case 16: LH_ELL_LAUNCH(16); break;
        case 32: LH_ELL_LAUNCH(32); break;
        default: LH_ELL_LAUNCH(64); break;
    }
#undef LH_ELL_LAUNCH
    LH_CHECK_LAUNCH();
    return true;
}
template bool spmm_ell<double, double>(lh_ctx*, const Ell&, const double*, double*, double*, int, int, int, int, int);
template bool spmm_ell<float, float>(lh_ctx*, const Ell&, const float*, float*, float*, int, int, int, int, int);

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
