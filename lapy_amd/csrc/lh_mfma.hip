// FP64 matrix-core (MFMA) kernels for the tall-skinny products of the block
// eigensolver: the Gram matrix C = U^T V and the update Y = alpha U C + beta Y,
// with U, V, Y row-major n x (p | q) blocks (n ~ 10^6, p, q <= 192).
//
// FP64 blocks use v_mfma_f64_16x16x4f64, FP32 blocks (the mixed-precision
// eigensolver) v_mfma_f32_16x16x4f32; both compute D[16x16] += A[16x4] B[4x16]
// per wave, lane l supplying A[l % 16][l / 16] and B[l / 16][l % 16].  Lane l
// holds D[l / 16 + 4 r][l % 16] (FP64) or D[4 (l / 16) + r][l % 16] (FP32) in
// accumulator r (r = 0..3).  FP32 products accumulate in FP32 within a wave's
// row range; the Gram partials are summed in FP64.
//
// Gram: the reduction runs over the rows.  A wave owns a 64 x 64 output tile
// and a range of rows; per group of 4 rows each lane loads 4 consecutive
// columns of U and of V (one 32-byte load each, 512 contiguous bytes per 16
// lanes), which feed the 4 x 4 MFMA blocks of the tile: block (bi, bj) uses
// the columns 4 i + bi of U and 4 j + bj of V.  That permutation of the tile's
// rows and columns is undone when the wave writes its partial tile.  The
// partials are summed in fixed order by a second kernel (deterministic).
// Narrow blocks (at most 9 useful blocks of 16 x 16) use a variant whose
// blocks hold contiguous columns, so the blocks beyond p or q are skipped.
//
// Update: a wave owns 64 rows x 64 columns of Y (64 x 16 for small n); per
// group of 16 columns of U (the summed index) each lane loads 4 consecutive
// values of its row (one 32-byte load), which serve as the A operand of 4
// successive MFMAs, and the matching 4 rows of C (L2-resident) as B.
#include "lh_dense.hpp"

#include <mutex>

namespace lh {

namespace {

typedef double d2_t __attribute__((ext_vector_type(2)));
typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

template <typename T>
struct Mf;
template <>
struct Mf<double> {
    typedef d4_t v4;
    static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // row of the 16 x 16 block held in accumulator r by the lanes of group kg
    static __device__ __forceinline__ int drow(int kg, int r) { return kg + 4 * r; }
};
template <>
struct Mf<float> {
    typedef f4_t v4;
    static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int drow(int kg, int r) { return 4 * kg + r; }
};

template <typename T>
__device__ __forceinline__ typename Mf<T>::v4 zero4() {
    return (typename Mf<T>::v4){T(0), T(0), T(0), T(0)};
}

// four consecutive values row[c .. c+3] of a row-major block, zero outside
// [0, ncols); VEC: the 4 values are 16-byte aligned (FP64: each pair) and in
// bounds as a whole
template <bool VEC, typename T>
__device__ __forceinline__ typename Mf<T>::v4 load4(const T* __restrict__ row, int c, int ncols) {
    typename Mf<T>::v4 v = zero4<T>();
    if (VEC && c + 3 < ncols) {
        if constexpr (sizeof(T) == 8) {
            const d2_t lo = *reinterpret_cast<const d2_t*>(row + c);
            const d2_t hi = *reinterpret_cast<const d2_t*>(row + c + 2);
            v[0] = lo[0]; v[1] = lo[1]; v[2] = hi[0]; v[3] = hi[1];
        } else {
            v = *reinterpret_cast<const f4_t*>(row + c);
        }
    } else {
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (c + q < ncols) v[q] = row[c + q];
    }
    return v;
}

// tile index -> (ti, tj); sym: upper triangle ti <= tj in row order
__device__ __forceinline__ void tile_of(int t, int tv, int sym, int& ti, int& tj) {
    if (!sym) { ti = t / tv; tj = t - ti * tv; return; }
    ti = 0;
    int rowlen = tv;
    while (t >= rowlen) { t -= rowlen; ++ti; --rowlen; }
    tj = ti + t;
}

template <bool VEC, typename T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
k_gram_mfma(int64_t n, int64_t rows_per_wave, int nwaves, const T* __restrict__ U, int ldu, int p,
            const T* __restrict__ V, int ldv, int q, int tv, int sym, double* __restrict__ part) {
    typedef typename Mf<T>::v4 v4;
    const int lane = threadIdx.x & 63;
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= nwaves) return;  // whole wave exits together
    const int tile = blockIdx.y;
    int ti, tj;
    tile_of(tile, tv, sym, ti, tj);
    const int u0 = ti * 64, v0 = tj * 64;
    const int il = lane & 15, kg = lane >> 4;
    const int64_t rb = (int64_t)w * rows_per_wave;
    const int64_t re = (rb + rows_per_wave < n) ? rb + rows_per_wave : n;
    v4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = zero4<T>();
    const int cu = u0 + 4 * il, cv = v0 + 4 * il;
    // 8 rows per step (2 MFMA k-groups of 4 rows); the next step's rows are
    // in flight during this step's 32 MFMAs
    auto load_rows = [&](int64_t r0, v4 (&a)[2], v4 (&b)[2]) {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int64_t r = r0 + 4 * s + kg;
            if (r < re) {
                a[s] = load4<VEC, T>(U + r * ldu, cu, p);
                b[s] = load4<VEC, T>(V + r * ldv, cv, q);
            } else {
                a[s] = zero4<T>();
                b[s] = zero4<T>();
            }
        }
    };
    v4 a[2], b[2];
    if (rb < re) load_rows(rb, a, b);
    for (int64_t r0 = rb; r0 < re; r0 += 8) {
        v4 an[2], bn[2];
        const bool more = r0 + 8 < re;
        if (more) load_rows(r0 + 8, an, bn);
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int bi = 0; bi < 4; ++bi)
#pragma unroll
                for (int bj = 0; bj < 4; ++bj) acc[bi][bj] = Mf<T>::mma(a[s][bi], b[s][bj], acc[bi][bj]);
        if (more) {
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                a[s] = an[s];
                b[s] = bn[s];
            }
        }
    }
    // partial tile [w][tile][64 x 64] in natural order
    double* dst = part + ((int64_t)w * gridDim.y + tile) * 4096;
#pragma unroll
    for (int bi = 0; bi < 4; ++bi)
#pragma unroll
        for (int bj = 0; bj < 4; ++bj)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = 4 * Mf<T>::drow(kg, r) + bi;  // tile-local column of U
                const int j = 4 * il + bj;                  // tile-local column of V
                dst[i * 64 + j] = (double)acc[bi][bj][r];
            }
}

// Gram of narrow blocks (p, q <= 64, at most 9 useful 16 x 16 blocks): block
// (bi, bj) holds the contiguous columns 16 bi .. 16 bi + 15 of U and 16 bj ..
// 16 bj + 15 of V, so the blocks wholly beyond p or q are skipped (a 16-column
// Gram runs one matrix-core block instead of 16).  Lane l loads U[r][16 bi +
// l % 16] of row r = 4 s + l / 16 (128 contiguous bytes per 16 lanes).  The
// partial tile has the layout of k_gram_mfma's; only its blocks bi < nbu,
// bj < nbv are written, which hold every entry i < p, j < q.
template <typename T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
k_gram_mfma_narrow(int64_t n, int64_t rows_per_wave, int nwaves, const T* __restrict__ U, int ldu, int p,
                   const T* __restrict__ V, int ldv, int q, double* __restrict__ part) {
    typedef typename Mf<T>::v4 v4;
    const int lane = threadIdx.x & 63;
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= nwaves) return;  // whole wave exits together
    const int il = lane & 15, kg = lane >> 4;
    const int nbu = (p + 15) >> 4, nbv = (q + 15) >> 4;
    const int64_t rb = (int64_t)w * rows_per_wave;
    const int64_t re = (rb + rows_per_wave < n) ? rb + rows_per_wave : n;
    v4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = zero4<T>();
    auto load_rows = [&](int64_t r0, T (&a)[2][4], T (&b)[2][4]) {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int64_t r = r0 + 4 * s + kg;
            const bool in = r < re;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int cc = 16 * c + il;
                a[s][c] = (c < nbu && in && cc < p) ? U[r * ldu + cc] : T(0);
                b[s][c] = (c < nbv && in && cc < q) ? V[r * ldv + cc] : T(0);
            }
        }
    };
    T a[2][4], b[2][4];
    if (rb < re) load_rows(rb, a, b);
    for (int64_t r0 = rb; r0 < re; r0 += 8) {
        T an[2][4], bn[2][4];
        const bool more = r0 + 8 < re;
        if (more) load_rows(r0 + 8, an, bn);
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int bi = 0; bi < 4; ++bi)
#pragma unroll
                for (int bj = 0; bj < 4; ++bj)
                    if (bi < nbu && bj < nbv) acc[bi][bj] = Mf<T>::mma(a[s][bi], b[s][bj], acc[bi][bj]);
        if (more) {
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    a[s][c] = an[s][c];
                    b[s][c] = bn[s][c];
                }
        }
    }
    double* dst = part + (int64_t)w * 4096;  // one tile
#pragma unroll
    for (int bi = 0; bi < 4; ++bi)
#pragma unroll
        for (int bj = 0; bj < 4; ++bj)
            if (bi < nbu && bj < nbv) {
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    dst[(16 * bi + Mf<T>::drow(kg, r)) * 64 + 16 * bj + il] = (double)acc[bi][bj][r];
            }
}

// C[i][j] = alpha * sum over waves (fixed order: 8 interleaved lane groups,
// combined pairwise).  One thread per element of a stored tile, so that
// consecutive lanes read consecutive partials; sym: the stored tiles are the
// upper ones (ti <= tj) and an off-diagonal tile also writes its mirror
// C[j][i].
__global__ void __launch_bounds__(512)
k_gram_mfma_reduce(int nwaves, int ntile, int p, int q, int tv, int sym, const double* __restrict__ part,
                   double* __restrict__ C, int ldc, double alpha) {
    __shared__ double sh[8][64];
    const int64_t o = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
    const int g = threadIdx.x >> 6;
    int i = p, j = q;
    int ti = 0, tj = 0;
    double s = 0.0;
    if (o < (int64_t)ntile * 4096) {
        const int tile = (int)(o >> 12), e = (int)(o & 4095);
        if (sym) {
            // tile = ti * tv - ti (ti - 1) / 2 + (tj - ti), ti <= tj
            int first = 0;
            while (first + (tv - ti) <= tile) {
                first += tv - ti;
                ++ti;
            }
            tj = ti + (tile - first);
        } else {
            ti = tile / tv;
            tj = tile - ti * tv;
        }
        i = ti * 64 + (e >> 6);
        j = tj * 64 + (e & 63);
        if (i < p && j < q) {
            const double* src = part + o;
            const int64_t stride = (int64_t)ntile * 4096;
            // eight loads in flight, summed in the order of c
            int c = g;
            for (; c + 56 < nwaves; c += 64) {
                double x[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) x[u] = src[(int64_t)(c + 8 * u) * stride];
#pragma unroll
                for (int u = 0; u < 8; ++u) s += x[u];
            }
            if (c + 24 < nwaves) {
                double x[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) x[u] = src[(int64_t)(c + 8 * u) * stride];
#pragma unroll
                for (int u = 0; u < 4; ++u) s += x[u];
                c += 32;
            }
            for (; c < nwaves; c += 8) s += src[c * stride];
        }
    }
    sh[g][threadIdx.x & 63] = s;
    __syncthreads();
    if (g == 0 && i < p && j < q) {
        const int l = threadIdx.x;
        const double v =
            alpha * (((sh[0][l] + sh[1][l]) + (sh[2][l] + sh[3][l])) + ((sh[4][l] + sh[5][l]) + (sh[6][l] + sh[7][l])));
        C[(int64_t)i * ldc + j] = v;
        if (sym && ti != tj) C[(int64_t)j * ldc + i] = v;
    }
}

// Y = alpha U C + beta Y.  upper: C is upper triangular (rows k > j of column
// j are zero), so a 64-column block only sums k < its last column + 1.
// GROUPED (q <= 256): a workgroup owns all column blocks of its rows, so the
// rows of U are fetched from HBM once for all of them.  INPLACE (Y = U,
// p = q <= 256, grouped): the workgroup writes only after every wave of it has
// read its rows of U.  A wave's tile is 64 rows x 16 NB columns: NB = 4 for
// large n; NB = 1 for small n, where 64-column tiles leave most of the GPU idle
// and each wave's chain of dependent matrix-core steps is four times shorter
// (U, small enough to stay in the caches, is then read once per 16 columns).
// nq counts the tile-wide column blocks.
template <bool VEC, bool INPLACE, bool GROUPED, int NB, typename T>
__global__ void __launch_bounds__(NB == 4 ? 256 : 512) __attribute__((amdgpu_waves_per_eu(2)))
k_gemm_mfma(int64_t n, int p, int q, const T* U, int ldu, const double* __restrict__ Cm, int ldc, T* Y, int ldy,
            double alpha, double beta, int upper, int nq) {
    constexpr int TW = 16 * NB;
    typedef typename Mf<T>::v4 v4;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int il = lane & 15, kg = lane >> 4;
    int64_t rw;  // first row of this wave
    int n0;      // first column of this wave
    if (INPLACE || GROUPED) {
        const int rsub = (int)(blockDim.x >> 6) / nq;
        rw = xcd_block() * 64 * rsub + (int64_t)(wv / nq) * 64;
        n0 = (wv % nq) * TW;
        if (!INPLACE && rw >= n) return;  // whole wave (no barrier follows)
    } else {
        rw = xcd_block() * 256 + (int64_t)wv * 64;
        n0 = blockIdx.y * TW;
        if (rw >= n) return;  // whole wave
    }
    const int kend = upper ? min(p, n0 + TW) : p;
    v4 acc[4][NB];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[a][b] = zero4<T>();
    // one K step's operands: 16 summed indices, this lane's 4 of them
    auto load_step = [&](int k0, v4 (&a)[4], T (&b)[4][NB]) {
        const int kc = k0 + 4 * kg;
#pragma unroll
        for (int bi = 0; bi < 4; ++bi) {
            const int64_t r = rw + 16 * bi + il;
            a[bi] = (r < n && k0 < kend) ? load4<VEC, T>(U + r * ldu, kc, kend) : zero4<T>();
        }
#pragma unroll
        for (int jj = 0; jj < 4; ++jj)
#pragma unroll
            for (int bj = 0; bj < NB; ++bj) {
                const int col = n0 + 16 * bj + il;
                b[jj][bj] = (kc + jj < kend && col < q) ? (T)Cm[(int64_t)(kc + jj) * ldc + col] : T(0);
            }
    };
    auto mma_step = [&](const v4 (&a)[4], const T (&b)[4][NB]) {
#pragma unroll
        for (int jj = 0; jj < 4; ++jj)
#pragma unroll
            for (int bi = 0; bi < 4; ++bi)
#pragma unroll
                for (int bj = 0; bj < NB; ++bj) acc[bi][bj] = Mf<T>::mma(a[bi][jj], b[jj][bj], acc[bi][bj]);
    };
    if constexpr (NB == 1) {
        // 16-column tiles serve small n, where a step's loads are not hidden
        // by other waves: the next step's operands are loaded before this
        // step's matrix-core products
        v4 a[4], an[4];
        T b[4][NB], bn[4][NB];
        load_step(0, a, b);
        for (int k0 = 0; k0 < kend; k0 += 16) {
            load_step(k0 + 16, an, bn);  // zeros past kend
            mma_step(a, b);
#pragma unroll
            for (int bi = 0; bi < 4; ++bi) a[bi] = an[bi];
#pragma unroll
            for (int jj = 0; jj < 4; ++jj)
#pragma unroll
                for (int bj = 0; bj < NB; ++bj) b[jj][bj] = bn[jj][bj];
        }
    } else {
        for (int k0 = 0; k0 < kend; k0 += 16) {
            v4 a[4];
            T b[4][NB];
            load_step(k0, a, b);
            mma_step(a, b);
        }
    }
    if (INPLACE) __syncthreads();
    // per row block: all of its Y values are read before any is written (a
    // load cannot pass a store the compiler cannot prove distinct, so
    // interleaving them would cost one memory round trip per value)
#pragma unroll
    for (int bi = 0; bi < 4; ++bi) {
        T yv[4][NB];
        if (beta != 0.0) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t row = rw + 16 * bi + Mf<T>::drow(kg, r);
#pragma unroll
                for (int bj = 0; bj < NB; ++bj) {
                    const int col = n0 + 16 * bj + il;
                    yv[r][bj] = (row < n && col < q) ? Y[row * ldy + col] : T(0);
                }
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t row = rw + 16 * bi + Mf<T>::drow(kg, r);
            if (row >= n) continue;
#pragma unroll
            for (int bj = 0; bj < NB; ++bj) {
                const int col = n0 + 16 * bj + il;
                if (col >= q) continue;
                double v = alpha * (double)acc[bi][bj][r];
                if (beta != 0.0) v += beta * (double)yv[r][bj];
                __builtin_nontemporal_store((T)v, Y + row * ldy + col);  // streamed out
            }
        }
    }
}

inline bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }
// 16-byte vector loads (FP64: pairs, FP32: quadruples) at column offsets
// divisible by 4
template <typename T>
bool vec_ok(const T* p, int ld) {
    return aligned16(p) && ld % (int)(16 / sizeof(T)) == 0;
}

}  // namespace

namespace {

// row split of a Gram: nw waves per output tile, rpw rows each (a multiple of
// 16); part holds nw * ntile partial tiles of 64 x 64 doubles
struct GramPlan {
    int tv = 0, ntile = 0;
    int64_t nw = 0, rpw = 0;
};

// about two waves per SIMD over all tiles, at least 64 rows per wave;
// max_parts > 0 caps nw * ntile (the fixed scratch of lh_block_gram_noalloc)
GramPlan gram_plan(const lh_ctx* ctx, int64_t n, int p, int q, bool sym, int64_t max_parts) {
    GramPlan g;
    const int tu = (p + 63) / 64;
    g.tv = (q + 63) / 64;
    g.ntile = sym ? tu * (tu + 1) / 2 : tu * g.tv;
    const int64_t target = (int64_t)ctx->n_cu * 8;
    int64_t nw = std::max<int64_t>(1, target / g.ntile);
    nw = std::min<int64_t>(nw, (n + 63) / 64);
    if (max_parts > 0) nw = std::min<int64_t>(nw, std::max<int64_t>(1, max_parts / g.ntile));
    int64_t rpw = (n + nw - 1) / nw;
    rpw = ((rpw + 15) / 16) * 16;
    g.rpw = rpw;
    g.nw = (n + rpw - 1) / rpw;  // <= the nw above
    return g;
}

// one output tile of which at most 9 of the 16 matrix-core blocks are needed
inline bool gram_narrow(int p, int q) { return p <= 64 && q <= 64 && ((p + 15) / 16) * ((q + 15) / 16) <= 9; }

template <typename T>
void gram_launch(lh_ctx* ctx, const GramPlan& g, int64_t n, const T* U, int ldu, int p, const T* V, int ldv, int q,
                 bool sym, double* part, double* C, int ldc, double alpha) {
    const bool vec = vec_ok(U, ldu) && vec_ok(V, ldv);
    dim3 grid((unsigned)((g.nw + 3) / 4), (unsigned)g.ntile);
    if (gram_narrow(p, q))
        k_gram_mfma_narrow<T><<<grid, 256, 0, ctx->stream>>>(n, g.rpw, (int)g.nw, U, ldu, p, V, ldv, q, part);
    else if (vec)
        k_gram_mfma<true, T><<<grid, 256, 0, ctx->stream>>>(n, g.rpw, (int)g.nw, U, ldu, p, V, ldv, q, g.tv,
                                                            sym ? 1 : 0, part);
    else
        k_gram_mfma<false, T><<<grid, 256, 0, ctx->stream>>>(n, g.rpw, (int)g.nw, U, ldu, p, V, ldv, q, g.tv,
                                                             sym ? 1 : 0, part);
    LH_CHECK_LAUNCH();
    k_gram_mfma_reduce<<<nblocks((int64_t)g.ntile * 4096, 64), 512, 0, ctx->stream>>>(
        (int)g.nw, g.ntile, p, q, g.tv, sym ? 1 : 0, part, C, ldc, alpha);
    LH_CHECK_LAUNCH();
}

}  // namespace

template <typename T>
void gram_mfma(lh_ctx* ctx, int64_t n, const T* U, int ldu, int p, const T* V, int ldv, int q, double* C, int ldc,
               DevBuf& ws, double alpha, bool sym) {
    if (sym && p != q) LH_THROW(LH_ERR_ARG, "symmetric Gram needs p == q");
    const GramPlan g = gram_plan(ctx, n, p, q, sym, 0);
    ws.ensure((size_t)g.nw * g.ntile * 4096 * sizeof(double));
    gram_launch<T>(ctx, g, n, U, ldu, p, V, ldv, q, sym, ws.as<double>(), C, ldc, alpha);
}

// 16-column tiles when 64-column ones would give fewer than 512 waves
inline bool gemm_narrow(int64_t n, int q) { return (n + 63) / 64 * ((q + 63) / 64) < 512; }

template <typename T>
void gemm_mfma(lh_ctx* ctx, int64_t n, int p, int q, const T* U, int ldu, const double* C, int ldc, T* Y, int ldy,
               double alpha, double beta, bool upper) {
    if (n == 0 || q == 0) return;
    const bool vec = vec_ok(U, ldu);
    const int up = upper ? 1 : 0;
    if (gemm_narrow(n, q)) {
        const int nq = (q + 15) / 16;
        if (nq > 1 && nq <= 8) {
            // grouped: nq 16-column blocks x rsub row blocks of 64 per
            // workgroup (at most 512 threads)
            const int rsub = std::max(1, 8 / nq);
            const int bs = 64 * nq * rsub;
            const unsigned nb = nblocks_xcd(n, 64 * rsub);
            if (vec)
                k_gemm_mfma<true, false, true, 1, T><<<nb, bs, 0, ctx->stream>>>(n, p, q, U, ldu, C, ldc, Y, ldy,
                                                                                alpha, beta, up, nq);
            else
                k_gemm_mfma<false, false, true, 1, T><<<nb, bs, 0, ctx->stream>>>(n, p, q, U, ldu, C, ldc, Y, ldy,
                                                                                 alpha, beta, up, nq);
        } else {
            dim3 grid(nblocks_xcd(n, 256), (unsigned)nq);
            if (vec)
                k_gemm_mfma<true, false, false, 1, T><<<grid, 256, 0, ctx->stream>>>(n, p, q, U, ldu, C, ldc, Y, ldy,
                                                                                    alpha, beta, up, nq);
            else
                k_gemm_mfma<false, false, false, 1, T><<<grid, 256, 0, ctx->stream>>>(n, p, q, U, ldu, C, ldc, Y,
                                                                                     ldy, alpha, beta, up, nq);
        }
        LH_CHECK_LAUNCH();
        return;
    }
    const int nq = (q + 63) / 64;
    if (nq > 1 && nq <= 4) {
        // grouped: nq column blocks x rsub row blocks of 64 per workgroup
        const int rsub = nq == 2 ? 2 : 1;
        const int bs = 64 * nq * rsub;
        const unsigned nb = nblocks_xcd(n, 64 * rsub);
        if (vec)
            k_gemm_mfma<true, false, true, 4, T><<<nb, bs, 0, ctx->stream>>>(n, p, q, U, ldu, C, ldc, Y, ldy, alpha,
                                                                            beta, up, nq);
        else
            k_gemm_mfma<false, false, true, 4, T><<<nb, bs, 0, ctx->stream>>>(n, p, q, U, ldu, C, ldc, Y, ldy, alpha,
                                                                             beta, up, nq);
        LH_CHECK_LAUNCH();
        return;
    }
    dim3 grid(nblocks_xcd(n, 256), (unsigned)nq);
    if (vec)
        k_gemm_mfma<true, false, false, 4, T><<<grid, 256, 0, ctx->stream>>>(n, p, q, U, ldu, C, ldc, Y, ldy, alpha,
                                                                             beta, up, nq);
    else
        k_gemm_mfma<false, false, false, 4, T><<<grid, 256, 0, ctx->stream>>>(n, p, q, U, ldu, C, ldc, Y, ldy, alpha,
                                                                              beta, up, nq);
    LH_CHECK_LAUNCH();
}

template <typename T>
void gemm_mfma_inplace(lh_ctx* ctx, int64_t n, int p, T* X, int ldx, const double* C, int ldc, double alpha,
                       bool upper) {
    if (n == 0 || p == 0) return;
    if (p > 256) LH_THROW(LH_ERR_ARG, "in-place block update needs width <= 256, got %d", p);
    const bool vec = vec_ok(X, ldx);
    const int up = upper ? 1 : 0;
    if (p <= 128 && gemm_narrow(n, p)) {
        // a workgroup covers all p columns of its rows: nq <= 8 waves of 16
        // columns, rsub row blocks of 64 (at most 512 threads)
        const int nq = (p + 15) / 16;
        const int rsub = std::max(1, 8 / nq);
        const int bs = 64 * nq * rsub;
        const unsigned nb = nblocks_xcd(n, 64 * rsub);
        if (vec)
            k_gemm_mfma<true, true, true, 1, T><<<nb, bs, 0, ctx->stream>>>(n, p, p, X, ldx, C, ldc, X, ldx, alpha,
                                                                           0.0, up, nq);
        else
            k_gemm_mfma<false, true, true, 1, T><<<nb, bs, 0, ctx->stream>>>(n, p, p, X, ldx, C, ldc, X, ldx, alpha,
                                                                            0.0, up, nq);
        LH_CHECK_LAUNCH();
        return;
    }
    const int nq = (p + 63) / 64;
    const int rsub = nq == 1 ? 4 : (nq == 2 ? 2 : 1);
    const int bs = 64 * nq * rsub;
    const unsigned nb = nblocks_xcd(n, 64 * rsub);
    if (vec)
        k_gemm_mfma<true, true, true, 4, T><<<nb, bs, 0, ctx->stream>>>(n, p, p, X, ldx, C, ldc, X, ldx, alpha, 0.0,
                                                                        up, nq);
    else
        k_gemm_mfma<false, true, true, 4, T><<<nb, bs, 0, ctx->stream>>>(n, p, p, X, ldx, C, ldc, X, ldx, alpha, 0.0,
                                                                         up, nq);
    LH_CHECK_LAUNCH();
}

template void gram_mfma<double>(lh_ctx*, int64_t, const double*, int, int, const double*, int, int, double*, int,
                                DevBuf&, double, bool);
template void gram_mfma<float>(lh_ctx*, int64_t, const float*, int, int, const float*, int, int, double*, int,
                               DevBuf&, double, bool);
template void gemm_mfma<double>(lh_ctx*, int64_t, int, int, const double*, int, const double*, int, double*, int,
                                double, double, bool);
template void gemm_mfma<float>(lh_ctx*, int64_t, int, int, const float*, int, const double*, int, float*, int, double,
                               double, bool);
template void gemm_mfma_inplace<double>(lh_ctx*, int64_t, int, double*, int, const double*, int, double, bool);
template void gemm_mfma_inplace<float>(lh_ctx*, int64_t, int, float*, int, const double*, int, double, bool);

// Scratch of lh_block_gram_noalloc: the per-wave partial tiles of its Gram, a
// fixed device array (each device holds its own copy while the process lives).
// Host code never passes the array's name to a kernel (on the host it is only a
// stand-in for the symbol): every use goes through the device address that
// hipGetSymbolAddress returned for the current device.
constexpr int64_t NOALLOC_PART_TILES = 512;  // 512 x 64 x 64 doubles = 16 MiB
static __device__ __align__(256) double scratch[NOALLOC_PART_TILES * 4096];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
}  // namespace

static double* noalloc_scratch(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return (device >= 0 && device < NOALLOC_MAX_DEVICES) ? g_scratch_ptr[device] : nullptr;
}

}  // namespace lh

using namespace lh;

void noalloc_scratch_init(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES)
        LH_THROW(LH_ERR_ARG, "device %d beyond the %d devices of the no-allocation scratch", device,
                 NOALLOC_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* p = nullptr;
    HIPX(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));  // the current device's copy
    g_scratch_ptr[device] = static_cast<double*>(p);
}

namespace {

// average duration (us) of `iters` launches of f on ctx's stream
template <typename F>
double time_launches(lh_ctx* ctx, int iters, F&& f) {
    hipEvent_t e0, e1;
    HIPX(hipEventCreate(&e0));
    HIPX(hipEventCreate(&e1));
    HIPX(hipEventRecord(e0, ctx->stream));
    for (int it = 0; it < iters; ++it) f();
    HIPX(hipEventRecord(e1, ctx->stream));
    HIPX(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPX(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 1000.0 * ms / iters;
}

// host FP64 block -> device block of T (converted on the device)
template <typename T>
void upload_block(lh_ctx* ctx, DevBuf& dst, const double* src, size_t count) {
    dst.ensure(count * sizeof(T));
    if constexpr (sizeof(T) == 8) {
        h2d(ctx, dst.p, src, count * 8);
    } else {
        DevBuf tmp(count * 8);
        h2d(ctx, tmp.p, src, count * 8);
        copy_block<double, T>(ctx, (int64_t)count, 1, tmp.as<double>(), 1, dst.as<T>(), 1);
    }
}

template <typename T>
void block_gram(lh_ctx* ctx, int64_t n, const double* U, int ldu, int p, const double* V, int ldv, int q, bool sym,
                double* C, int iters, double* avg_us) {
    DevBuf du, dv, dc((size_t)p * q * 8), ws;
    upload_block<T>(ctx, du, U, (size_t)n * ldu);
    upload_block<T>(ctx, dv, V, (size_t)n * ldv);
    auto run = [&] { gram_mfma<T>(ctx, n, du.as<T>(), ldu, p, dv.as<T>(), ldv, q, dc.as<double>(), q, ws, 1.0, sym); };
    run();
    if (avg_us && iters > 0) *avg_us = time_launches(ctx, iters, run);
    d2h(ctx, C, dc.p, (size_t)p * q * 8);
}

template <typename T>
void block_gemm(lh_ctx* ctx, int64_t n, int p, int q, const double* U, int ldu, const double* Cs, double* Y, int ldy,
                double alpha, double beta, bool upper, bool inplace, int iters, double* avg_us) {
    DevBuf du, dy, dc((size_t)p * q * 8);
    if (!inplace) upload_block<T>(ctx, du, U, (size_t)n * ldu);
    upload_block<T>(ctx, dy, Y, (size_t)n * ldy);
    h2d(ctx, dc.p, Cs, (size_t)p * q * 8);
    auto run = [&](double b) {
        if (inplace) gemm_mfma_inplace<T>(ctx, n, p, dy.as<T>(), ldy, dc.as<double>(), q, alpha, upper);
        else gemm_mfma<T>(ctx, n, p, q, du.as<T>(), ldu, dc.as<double>(), q, dy.as<T>(), ldy, alpha, b, upper);
    };
    run(beta);
    if constexpr (sizeof(T) == 8) {
        d2h(ctx, Y, dy.p, (size_t)n * ldy * 8);
    } else {
        DevBuf tmp((size_t)n * ldy * 8);
        copy_block<T, double>(ctx, n * ldy, 1, dy.as<T>(), 1, tmp.as<double>(), 1);
        d2h(ctx, Y, tmp.p, (size_t)n * ldy * 8);
    }
    if (avg_us && iters > 0) *avg_us = time_launches(ctx, iters, [&] { run(0.0); });
}

}  // namespace

extern "C" {

int lh_block_gram(lh_ctx* ctx, int64_t n, const double* U, int ldu, int p, const double* V, int ldv, int q, int flags,
                  double* C, int iters, double* avg_us) {
    if (!ctx || !U || !V || !C || n <= 0 || p <= 0 || q <= 0 || ldu < p || ldv < q) return LH_ERR_ARG;
    if (p > 1024 || q > 1024) return LH_ERR_ARG;
    const bool sym = (flags & LH_GRAM_SYM) != 0, f32 = (flags & LH_BLOCK_F32) != 0;
    LH_API_BEGIN(ctx)
    HIPX(hipSetDevice(ctx->device));
    if (sym && p != q) LH_THROW(LH_ERR_ARG, "symmetric Gram needs p == q");
    if (f32) block_gram<float>(ctx, n, U, ldu, p, V, ldv, q, sym, C, iters, avg_us);
    else block_gram<double>(ctx, n, U, ldu, p, V, ldv, q, sym, C, iters, avg_us);
    LH_API_END(ctx)
}

int lh_block_gram_noalloc(lh_ctx* ctx, int64_t n, const double* U, int ldu, int p, const double* V, int ldv, int q,
                          int flags, double* C) {
    if (!ctx || !U || !V || !C || n <= 0 || p <= 0 || q <= 0 || ldu < p || ldv < q) return LH_ERR_ARG;
    LH_API_BEGIN(ctx)
    if (flags & ~LH_GRAM_SYM) LH_THROW(LH_ERR_ARG, "lh_block_gram_noalloc: FP64 blocks only (flags %d)", flags);
    if (p > LH_GRAM_NOALLOC_MAX_COLS || q > LH_GRAM_NOALLOC_MAX_COLS)
        LH_THROW(LH_ERR_ARG, "lh_block_gram_noalloc: %d x %d Gram exceeds its fixed scratch (at most %d columns)", p,
                 q, LH_GRAM_NOALLOC_MAX_COLS);
    const bool sym = (flags & LH_GRAM_SYM) != 0;
    if (sym && p != q) LH_THROW(LH_ERR_ARG, "symmetric Gram needs p == q");
    HIPX(hipSetDevice(ctx->device));
    double* part = noalloc_scratch(ctx->device);
    if (!part) LH_THROW(LH_ERR_HIP, "no-allocation scratch of device %d not set up", ctx->device);
    const GramPlan g = gram_plan(ctx, n, p, q, sym, NOALLOC_PART_TILES);
    if (g.nw * g.ntile > NOALLOC_PART_TILES)
        LH_THROW(LH_ERR_ARG, "lh_block_gram_noalloc: %lld partial tiles exceed the scratch's %lld",
                 (long long)(g.nw * g.ntile), (long long)NOALLOC_PART_TILES);
    gram_launch<double>(ctx, g, n, U, ldu, p, V, ldv, q, sym, part, C, q, 1.0);
    LH_API_END(ctx)
}

int lh_block_gemm(lh_ctx* ctx, int64_t n, int p, int q, const double* U, int ldu, const double* Cs, double* Y,
                  int ldy, double alpha, double beta, int flags, int iters, double* avg_us) {
    if (!ctx || !Cs || !Y || n <= 0 || p <= 0 || q <= 0 || ldy < q) return LH_ERR_ARG;
    if (p > 1024 || q > 1024) return LH_ERR_ARG;
    const bool upper = (flags & LH_GEMM_UPPER) != 0, inplace = (flags & LH_GEMM_INPLACE) != 0;
    const bool f32 = (flags & LH_BLOCK_F32) != 0;
    if (inplace && (p != q || p > 256 || beta != 0.0)) return LH_ERR_ARG;
    if (!inplace && (!U || ldu < p)) return LH_ERR_ARG;
    LH_API_BEGIN(ctx)
    HIPX(hipSetDevice(ctx->device));
    if (f32) block_gemm<float>(ctx, n, p, q, U, ldu, Cs, Y, ldy, alpha, beta, upper, inplace, iters, avg_us);
    else block_gemm<double>(ctx, n, p, q, U, ldu, Cs, Y, ldy, alpha, beta, upper, inplace, iters, avg_us);
    LH_API_END(ctx)
}

}  // extern "C"
