// Locality-improving symmetric reordering of a CSR matrix, computed on the GPU
// and applied inside the iterative solvers.
//
// Mesh vertex numberings are often far from local (1-to-4 subdivision
// appends every edge midpoint at the end, so most edges join vertices far
// apart in index).  The SpMM-type kernels gather one row of a block per
// nonzero; in a local order the rows a workgroup gathers lie close to its
// own, so the gathers hit the XCD's L2 instead of HBM.  A mesh's pattern
// carries the Morton codes of its vertices (30 bits on a 1024^3 grid over the
// bounding box, computed with the pattern): the order sorts them, one radix
// sort.  A pattern without coordinates (a matrix handed in by the caller)
// gets breadth-first levels instead, one launch per level.
//
// Deterministic: the permutation is a stable sort of (key, original index),
// key = Morton code or BFS level (a graph distance), with an LSD radix sort
// whose ranks are computed per workgroup without atomics.  Results are mapped
// back to the caller's numbering, so the permutation is invisible outside the
// solver.
#include "lh_dense.hpp"

namespace lh {

namespace {

__global__ void k_fill_i32(int32_t* __restrict__ x, int64_t n, int32_t v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

__global__ void k_bfs_step(int64_t n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cols,
                           int32_t* __restrict__ level, int32_t cur, int32_t* __restrict__ changed) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || level[i] != cur) return;
    bool any = false;
    for (int32_t s = rowptr[i]; s < rowptr[i + 1]; ++s) {
        const int32_t j = cols[s];
        if (level[j] < 0) {  // every writer stores the same value
            level[j] = cur + 1;
            any = true;
        }
    }
    if (any) changed[0] = 1;
}

__global__ void k_unreached(int64_t n, int32_t* __restrict__ level, int32_t last) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && level[i] < 0) level[i] = last;
}

__global__ void k_iota(int32_t* __restrict__ x, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = (int32_t)i;
}

constexpr int RBS = 256;

// per-block histogram of an 8-bit digit: hist[digit * nb + block]
__global__ void __launch_bounds__(RBS)
k_digit_hist(const int32_t* __restrict__ keys, int64_t n, int shift, int32_t* __restrict__ hist, int nb) {
    __shared__ int32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * RBS + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255], 1);
    __syncthreads();
    hist[(int64_t)threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

// stable scatter: rank of an element among earlier same-digit elements of its block
__global__ void __launch_bounds__(RBS)
k_digit_scatter(const int32_t* __restrict__ kin, const int32_t* __restrict__ vin, int64_t n, int shift,
                const int32_t* __restrict__ off, int nb, int32_t* __restrict__ kout, int32_t* __restrict__ vout) {
    __shared__ int32_t dg[RBS];
    const int64_t i = (int64_t)blockIdx.x * RBS + threadIdx.x;
    const int d = (i < n) ? ((kin[i] >> shift) & 255) : -1;
    dg[threadIdx.x] = d;
    __syncthreads();
    if (i >= n) return;
    int rank = 0;
    for (int t = 0; t < (int)threadIdx.x; ++t) rank += (dg[t] == d);
    const int32_t pos = off[(int64_t)d * nb + blockIdx.x] + rank;
    kout[pos] = kin[i];
    vout[pos] = vin[i];
}

__global__ void k_invert(const int32_t* __restrict__ perm, int32_t* __restrict__ iperm, int64_t n) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) iperm[perm[r]] = (int32_t)r;
}

__global__ void k_perm_rowlen(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ perm,
                              int32_t* __restrict__ len, int64_t n) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) len[r] = rowptr[perm[r] + 1] - rowptr[perm[r]];
}

// new row r: columns iperm[old cols] sorted, carrying the source slot
__global__ void k_perm_rows(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cols,
                            const int32_t* __restrict__ perm, const int32_t* __restrict__ iperm,
                            const int32_t* __restrict__ nrowptr, int32_t* __restrict__ ncols,
                            int32_t* __restrict__ src, int32_t* __restrict__ ndiag, int64_t n) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int32_t o = perm[r];
    const int32_t a = rowptr[o], b = rowptr[o + 1], base = nrowptr[r];
    for (int32_t s = a; s < b; ++s) {
        const int32_t c = iperm[cols[s]];
        int32_t q = base + (s - a) - 1;
        while (q >= base && ncols[q] > c) {  // insertion sort as we go
            ncols[q + 1] = ncols[q];
            src[q + 1] = src[q];
            --q;
        }
        ncols[q + 1] = c;
        src[q + 1] = s;
    }
    int32_t dp = -1;
    for (int32_t q = base; q < base + (b - a); ++q)
        if (ncols[q] == (int32_t)r) dp = q;
    ndiag[r] = dp;
}

template <typename T>
__global__ void k_gather_vals(const T* __restrict__ v, const int32_t* __restrict__ src, T* __restrict__ out,
                              int64_t nnz) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nnz) out[s] = v[src[s]];
}

template <typename T>
__global__ void k_gather_rows(const T* __restrict__ v, const int32_t* __restrict__ perm, T* __restrict__ out,
                              int64_t n) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) out[r] = v[perm[r]];
}

}  // namespace

// breadth-first levels of the pattern's graph into level (n entries): from
// vertex 0, then again from the first vertex of the deepest level (a
// pseudo-peripheral start gives narrower levels); unreached vertices get the
// level after the deepest.  Returns the number of levels.
static int32_t bfs_levels(lh_ctx* ctx, const Pattern& p, DevBuf& level) {
    const int64_t n = p.n;
    const int32_t* rp = p.rowptr.as<int32_t>();
    const int32_t* cl = p.cols.as<int32_t>();
    DevBuf changed(64);
    int32_t last = 0;
    int32_t start = 0;
    for (int round = 0; round < 2; ++round) {
        k_fill_i32<<<nblocks(n, 256), 256, 0, ctx->stream>>>(level.as<int32_t>(), n, -1);
        k_fill_i32<<<1, 1, 0, ctx->stream>>>(level.as<int32_t>() + start, 1, 0);
        LH_CHECK_LAUNCH();
        int32_t cur = 0;
        while (true) {
            // a few levels per host check
            k_fill_i32<<<1, 8, 0, ctx->stream>>>(changed.as<int32_t>(), 8, 0);
            for (int q = 0; q < 8; ++q)
                k_bfs_step<<<nblocks(n, 256), 256, 0, ctx->stream>>>(n, rp, cl, level.as<int32_t>(), cur + q,
                                                                    changed.as<int32_t>() + q);
            LH_CHECK_LAUNCH();
            int32_t h[8];
            d2h(ctx, h, changed.p, sizeof(h));
            int q = 0;
            while (q < 8 && h[q]) ++q;
            cur += q;
            if (q < 8) break;
            if (cur >= (1 << 24) - 16) LH_THROW(LH_ERR_NUMERIC, "reorder: too many BFS levels");
        }
        last = cur;  // deepest level reached
        if (round == 0) {
            // first vertex (lowest index) on the deepest level
            std::vector<int32_t> hl(n);
            d2h(ctx, hl.data(), level.p, n * 4);
            start = 0;
            for (int64_t i = 0; i < n; ++i)
                if (hl[i] == last) { start = (int32_t)i; break; }
        }
    }
    k_unreached<<<nblocks(n, 256), 256, 0, ctx->stream>>>(n, level.as<int32_t>(), last + 1);
    LH_CHECK_LAUNCH();
    return last + 2;
}

std::shared_ptr<Reorder> pattern_reorder(lh_ctx* ctx, Pattern& p) {
    std::lock_guard<std::mutex> lk(p.ro_mu);
    if (!p.ro) {
        ArenaOff off;  // kept with the pattern, beyond this call
        auto r = std::make_shared<Reorder>();
        r->build(ctx, p);
        p.ro = r;
    }
    return p.ro;
}

void Reorder::build(lh_ctx* ctx, const Pattern& p) {
    const int64_t n = p.n;
    const int32_t* rp = p.rowptr.as<int32_t>();
    const int32_t* cl = p.cols.as<int32_t>();
    DevBuf level(n * 4);
    int key_bits = 24;
    if (p.order_key.p) {
        // a mesh's Morton codes: sort by them (30 bits) instead of a search
        HIPX(hipMemcpyAsync(level.p, p.order_key.p, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
        key_bits = 32;
        nlevels = 0;
    } else {
        nlevels = bfs_levels(ctx, p, level);
    }
    // stable LSD radix sort of (key, index) by key, 8-bit digits
    const int nb = (int)((n + RBS - 1) / RBS);
    DevBuf k2(n * 4), v1(n * 4), v2(n * 4), hist((size_t)256 * nb * 4), off(((size_t)256 * nb + 1) * 4);
    k_iota<<<nblocks(n, 256), 256, 0, ctx->stream>>>(v1.as<int32_t>(), n);
    int32_t *kin = level.as<int32_t>(), *kout = k2.as<int32_t>(), *vin = v1.as<int32_t>(), *vout = v2.as<int32_t>();
    for (int shift = 0; shift < key_bits; shift += 8) {
        k_digit_hist<<<nb, RBS, 0, ctx->stream>>>(kin, n, shift, hist.as<int32_t>(), nb);
        LH_CHECK_LAUNCH();
        exclusive_scan_i32(ctx, hist.as<int32_t>(), off.as<int32_t>(), (int64_t)256 * nb);
        k_digit_scatter<<<nb, RBS, 0, ctx->stream>>>(kin, vin, n, shift, off.as<int32_t>(), nb, kout, vout);
        LH_CHECK_LAUNCH();
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    perm.alloc(n * 4);
    iperm.alloc(n * 4);
    HIPX(hipMemcpyAsync(perm.p, vin, n * 4, hipMemcpyDeviceToDevice, ctx->stream));
    k_invert<<<nblocks(n, 256), 256, 0, ctx->stream>>>(perm.as<int32_t>(), iperm.as<int32_t>(), n);
    LH_CHECK_LAUNCH();
    // permuted pattern
    pat = std::make_shared<Pattern>();
    pat->n = n;
    pat->nnz = p.nnz;
    DevBuf len(n * 4);
    k_perm_rowlen<<<nblocks(n, 256), 256, 0, ctx->stream>>>(rp, perm.as<int32_t>(), len.as<int32_t>(), n);
    LH_CHECK_LAUNCH();
    pat->rowptr.alloc((n + 1) * 4);
    exclusive_scan_i32(ctx, len.as<int32_t>(), pat->rowptr.as<int32_t>(), n);
    pat->cols.alloc(p.nnz * 4);
    pat->diagpos.alloc(n * 4);
    src.alloc(p.nnz * 4);
    k_perm_rows<<<nblocks(n, 128), 128, 0, ctx->stream>>>(rp, cl, perm.as<int32_t>(), iperm.as<int32_t>(),
                                                         pat->rowptr.as<int32_t>(), pat->cols.as<int32_t>(),
                                                         src.as<int32_t>(), pat->diagpos.as<int32_t>(), n);
    LH_CHECK_LAUNCH();
    HIPX(hipStreamSynchronize(ctx->stream));
}

std::unique_ptr<lh_csr> Reorder::apply(lh_ctx* ctx, const lh_csr* a, std::shared_ptr<Pattern> for_pattern,
                                       const Reorder* diag_reorder) const {
    auto out = std::make_unique<lh_csr>();
    out->ctx = ctx;
    out->dtype = a->dtype;
    out->is_diag = a->is_diag;
    const size_t es = a->dtype == LH_F64 ? 8 : 4;
    if (a->is_diag) {
        // diagonal matrix: permute the dense diagonal, pattern = full diagonal
        (void)diag_reorder;
        const int64_t n = a->n();
        out->pat = std::make_shared<Pattern>();
        out->pat->n = n;
        out->pat->nnz = n;
        std::vector<int32_t> rp(n + 1), cl(n);
        for (int64_t i = 0; i <= n; ++i) rp[i] = (int32_t)i;
        for (int64_t i = 0; i < n; ++i) cl[i] = (int32_t)i;
        out->pat->rowptr.alloc((n + 1) * 4);
        out->pat->cols.alloc(n * 4);
        out->pat->diagpos.alloc(n * 4);
        h2d(ctx, out->pat->rowptr.p, rp.data(), (n + 1) * 4);
        h2d(ctx, out->pat->cols.p, cl.data(), n * 4);
        h2d(ctx, out->pat->diagpos.p, cl.data(), n * 4);
        out->vals.alloc(n * es);
        out->dense_diag.alloc(n * es);
        if (a->dtype == LH_F64) {
            k_gather_rows<double><<<nblocks(n, 256), 256, 0, ctx->stream>>>(a->dense_diag.as<double>(),
                                                                           perm.as<int32_t>(),
                                                                           out->dense_diag.as<double>(), n);
        } else {
            k_gather_rows<float><<<nblocks(n, 256), 256, 0, ctx->stream>>>(a->dense_diag.as<float>(),
                                                                          perm.as<int32_t>(),
                                                                          out->dense_diag.as<float>(), n);
        }
        LH_CHECK_LAUNCH();
        HIPX(hipMemcpyAsync(out->vals.p, out->dense_diag.p, n * es, hipMemcpyDeviceToDevice, ctx->stream));
        return out;
    }
    out->pat = for_pattern ? for_pattern : pat;
    const int64_t nnz = a->nnz();
    out->vals.alloc(nnz * es);
    if (a->dtype == LH_F64)
        k_gather_vals<double><<<nblocks(nnz, 256), 256, 0, ctx->stream>>>(a->vals.as<double>(), src.as<int32_t>(),
                                                                         out->vals.as<double>(), nnz);
    else
        k_gather_vals<float><<<nblocks(nnz, 256), 256, 0, ctx->stream>>>(a->vals.as<float>(), src.as<int32_t>(),
                                                                        out->vals.as<float>(), nnz);
    LH_CHECK_LAUNCH();
    return out;
}

// out[old i, c] = in[new iperm[i], c]  for row-major n x m blocks
__global__ void k_unpermute_block(const double* __restrict__ in, int ldi, const int32_t* __restrict__ iperm,
                                  double* __restrict__ out, int ldo, int64_t n, int m) {
    int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n * m) return;
    int64_t i = gid / m;
    int c = (int)(gid - i * m);
    out[i * ldo + c] = in[(int64_t)iperm[i] * ldi + c];
}

void Reorder::unpermute(lh_ctx* ctx, const double* in, int ldi, double* out, int ldo, int64_t n, int m) const {
    k_unpermute_block<<<nblocks(n * m, 256), 256, 0, ctx->stream>>>(in, ldi, iperm.as<int32_t>(), out, ldo, n, m);
    LH_CHECK_LAUNCH();
}

}  // namespace lh
