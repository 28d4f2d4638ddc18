// Context, CSR handles, scans and copies.
#include "lh_internal.hpp"

#include <cstring>

namespace lh {

// ---------------------------------------------------------------------------
// exclusive scan (int32), three phases: per-block scan, scan of block sums
// (recursive), add block offsets.  Deterministic (integer).
// ---------------------------------------------------------------------------
constexpr int SCAN_BS = 256;
constexpr int SCAN_PER = 8;             // elements per thread
constexpr int SCAN_TILE = SCAN_BS * SCAN_PER;

__global__ void __launch_bounds__(SCAN_BS)
k_scan_tile(const int32_t* __restrict__ in, int32_t* __restrict__ out, int64_t n,
            int32_t* __restrict__ block_sums) {
    __shared__ int32_t s[SCAN_BS];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_PER;
    int32_t v[SCAN_PER];
    int32_t sum = 0;
#pragma unroll
    for (int q = 0; q < SCAN_PER; ++q) {
        int64_t i = base + q;
        v[q] = (i < n) ? in[i] : 0;
        sum += v[q];
    }
    s[threadIdx.x] = sum;
    __syncthreads();
    // Hillis-Steele inclusive scan over thread sums
    for (int off = 1; off < SCAN_BS; off <<= 1) {
        int32_t add = (threadIdx.x >= (unsigned)off) ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += add;
        __syncthreads();
    }
    int32_t run = s[threadIdx.x] - sum;   // exclusive prefix of this thread
#pragma unroll
    for (int q = 0; q < SCAN_PER; ++q) {
        int64_t i = base + q;
        if (i < n) out[i] = run;
        run += v[q];
    }
    if (threadIdx.x == SCAN_BS - 1) block_sums[blockIdx.x] = s[SCAN_BS - 1];
}

__global__ void k_scan_add(int32_t* __restrict__ out, int64_t n, const int32_t* __restrict__ offs) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += offs[i / SCAN_TILE];
}

__global__ void k_scan_total(int32_t* out, int64_t n, const int32_t* in) {
    // out[n] = out[n-1] + in[n-1]
    if (n > 0) out[n] = out[n - 1] + in[n - 1];
    else out[0] = 0;
}

static void scan_rec(lh_ctx* ctx, const int32_t* in, int32_t* out, int64_t n) {
    const int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    DevBuf sums((size_t)(nb + 1) * sizeof(int32_t));
    DevBuf sums_scan((size_t)(nb + 1) * sizeof(int32_t));
    k_scan_tile<<<(unsigned)nb, SCAN_BS, 0, ctx->stream>>>(in, out, n, sums.as<int32_t>());
    LH_CHECK_LAUNCH();
    if (nb > 1) {
        scan_rec(ctx, sums.as<int32_t>(), sums_scan.as<int32_t>(), nb);
        k_scan_add<<<nblocks(n, 256), 256, 0, ctx->stream>>>(out, n, sums_scan.as<int32_t>());
        LH_CHECK_LAUNCH();
    }
    HIPX(hipStreamSynchronize(ctx->stream));  // sums buffers go out of scope
}

void exclusive_scan_i32(lh_ctx* ctx, const int32_t* in, int32_t* out, int64_t n) {
    if (n > 0) scan_rec(ctx, in, out, n);
    k_scan_total<<<1, 1, 0, ctx->stream>>>(out, n, in);
    LH_CHECK_LAUNCH();
}

// Small transfers go through the context's pinned staging buffer: a copy
// between device and pinned memory is one DMA (or kernel) on the stream,
// where pageable memory takes the runtime's shared staging path, which
// serialises the concurrent solver streams of the batched eigensolver.
void d2h_n(lh_ctx* ctx, const Xfer* x, int cnt) {
    size_t total = 0;
    for (int i = 0; i < cnt; ++i) total += (x[i].bytes + 63) & ~size_t(63);
    if (!ctx->host_stage || total > ctx->host_stage_bytes) {
        for (int i = 0; i < cnt; ++i)
            if (x[i].bytes) HIPX(hipMemcpyAsync(x[i].host, x[i].dev, x[i].bytes, hipMemcpyDeviceToHost, ctx->stream));
        HIPX(hipStreamSynchronize(ctx->stream));
        return;
    }
    char* stage = static_cast<char*>(ctx->host_stage);
    size_t off = 0;
    for (int i = 0; i < cnt; ++i) {
        if (x[i].bytes) HIPX(hipMemcpyAsync(stage + off, x[i].dev, x[i].bytes, hipMemcpyDeviceToHost, ctx->stream));
        off += (x[i].bytes + 63) & ~size_t(63);
    }
    HIPX(hipStreamSynchronize(ctx->stream));
    off = 0;
    for (int i = 0; i < cnt; ++i) {
        if (x[i].bytes) std::memcpy(x[i].host, stage + off, x[i].bytes);
        off += (x[i].bytes + 63) & ~size_t(63);
    }
}
void d2h(lh_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!bytes) return;
    const Xfer x{dst, src, bytes};
    d2h_n(ctx, &x, 1);
}
void h2d(lh_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!bytes) return;
    if (ctx->host_stage && bytes <= ctx->host_stage_bytes) {
        std::memcpy(ctx->host_stage, src, bytes);
        HIPX(hipMemcpyAsync(dst, ctx->host_stage, bytes, hipMemcpyHostToDevice, ctx->stream));
    } else {
        HIPX(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    }
    HIPX(hipStreamSynchronize(ctx->stream));  // the source may be reused on return
}
void memset_d(lh_ctx* ctx, void* dst, int v, size_t bytes) {
    if (!bytes) return;
    HIPX(hipMemsetAsync(dst, v, bytes, ctx->stream));
}

// diagonal pattern restricted to rows flagged in used_flag
__global__ void k_flag_to_count(const int32_t* __restrict__ used, int32_t* __restrict__ cnt, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) cnt[i] = used[i] ? 1 : 0;
}
__global__ void k_diag_fill(const int32_t* __restrict__ rowptr, int32_t* __restrict__ cols,
                            int32_t* __restrict__ diagpos, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t r0 = rowptr[i], r1 = rowptr[i + 1];
    if (r1 > r0) { cols[r0] = (int32_t)i; diagpos[i] = r0; }
    else diagpos[i] = -1;
}

std::shared_ptr<Pattern> make_diag_pattern(lh_ctx* ctx, const Pattern& /*full*/, int64_t n,
                                           const int32_t* used_flag) {
    auto p = std::make_shared<Pattern>();
    p->n = n;
    p->rowptr.alloc((n + 1) * sizeof(int32_t));
    DevBuf cnt(n * sizeof(int32_t));
    k_flag_to_count<<<nblocks(n, 256), 256, 0, ctx->stream>>>(used_flag, cnt.as<int32_t>(), n);
    LH_CHECK_LAUNCH();
    exclusive_scan_i32(ctx, cnt.as<int32_t>(), p->rowptr.as<int32_t>(), n);
    int32_t nnz = 0;
    d2h(ctx, &nnz, p->rowptr.as<int32_t>() + n, sizeof(int32_t));
    p->nnz = nnz;
    p->cols.alloc((size_t)nnz * sizeof(int32_t));
    p->diagpos.alloc(n * sizeof(int32_t));
    k_diag_fill<<<nblocks(n, 256), 256, 0, ctx->stream>>>(p->rowptr.as<int32_t>(), p->cols.as<int32_t>(),
                                                         p->diagpos.as<int32_t>(), n);
    LH_CHECK_LAUNCH();
    HIPX(hipStreamSynchronize(ctx->stream));
    return p;
}

__global__ void k_diagpos(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cols,
                          int32_t* __restrict__ diagpos, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t d = -1;
    for (int32_t s = rowptr[i]; s < rowptr[i + 1]; ++s)
        if (cols[s] == (int32_t)i) { d = s; break; }
    diagpos[i] = d;
}

template <typename T>
__global__ void k_gather_diag(const T* __restrict__ vals, const int32_t* __restrict__ diagpos,
                              T* __restrict__ dd, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t d = diagpos[i];
    dd[i] = d >= 0 ? vals[d] : T(0);
}

void refresh_dense_diag(lh_ctx* ctx, lh_csr* a) {
    if (!a->is_diag) return;
    const int64_t n = a->n();
    size_t es = a->dtype == LH_F64 ? 8 : 4;
    a->dense_diag.ensure(n * es);
    if (a->dtype == LH_F64)
        k_gather_diag<double><<<nblocks(n, 256), 256, 0, ctx->stream>>>(
            a->vals.as<double>(), a->pat->diagpos.as<int32_t>(), a->dense_diag.as<double>(), n);
    else
        k_gather_diag<float><<<nblocks(n, 256), 256, 0, ctx->stream>>>(
            a->vals.as<float>(), a->pat->diagpos.as<int32_t>(), a->dense_diag.as<float>(), n);
    LH_CHECK_LAUNCH();
}

}  // namespace lh

using namespace lh;

thread_local Arena* lh_thread_arena = nullptr;

extern "C" {

int lh_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int lh_ctx_create(int device, lh_ctx** out) {
    if (!out) return LH_ERR_ARG;
    *out = nullptr;
    lh_ctx* ctx = new lh_ctx();
    ctx->scratch.persistent = ctx->small.persistent = ctx->pinned_dummy.persistent = true;
    ctx->arena_mem.persistent = true;
    try {
        int nd = 0;
        HIPX(hipGetDeviceCount(&nd));
        if (device < 0 || device >= nd) LH_THROW(LH_ERR_ARG, "device %d out of range (%d devices)", device, nd);
        ctx->device = device;
        HIPX(hipSetDevice(device));
        HIPX(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        hipDeviceProp_t prop;
        HIPX(hipGetDeviceProperties(&prop, device));
        ctx->n_cu = prop.multiProcessorCount;
        noalloc_scratch_init(device);
        ctx->host_small_bytes = 1 << 20;
        HIPX(hipHostMalloc(&ctx->host_small, ctx->host_small_bytes, hipHostMallocDefault));
        ctx->host_stage_bytes = 256 << 10;
        HIPX(hipHostMalloc(&ctx->host_stage, ctx->host_stage_bytes, hipHostMallocDefault));
    } catch (const lh_error& e) {
        fprintf(stderr, "lh_ctx_create: %s\n", e.msg.c_str());
        if (ctx->host_small) (void)hipHostFree(ctx->host_small);
        if (ctx->host_stage) (void)hipHostFree(ctx->host_stage);
        if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return e.code;
    }
    *out = ctx;
    return LH_OK;
}

void lh_ctx_destroy(lh_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (lh_ctx* w : ctx->workers) lh_ctx_destroy(w);
    ctx->workers.clear();
    ctx->arena_mem.release();
    ctx->scratch.release();
    ctx->small.release();
    if (ctx->host_small) (void)hipHostFree(ctx->host_small);
    if (ctx->host_stage) (void)hipHostFree(ctx->host_stage);
    release_blas(ctx);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* lh_last_error(lh_ctx* ctx) { return ctx ? ctx->last_error.c_str() : "no context"; }

int lh_sync(lh_ctx* ctx) {
    LH_API_BEGIN(ctx)
    HIPX(hipStreamSynchronize(ctx->stream));
    LH_API_END(ctx)
}

int lh_csr_info(lh_csr* a, int64_t* n, int64_t* nnz, int* dtype, int* is_diag) {
    if (!a) return LH_ERR_ARG;
    if (n) *n = a->n();
    if (nnz) *nnz = a->nnz();
    if (dtype) *dtype = a->dtype;
    if (is_diag) *is_diag = a->is_diag;
    return LH_OK;
}

int lh_csr_download(lh_csr* a, int32_t* indptr, int32_t* indices, void* data) {
    if (!a) return LH_ERR_ARG;
    lh_ctx* ctx = a->ctx;
    LH_API_BEGIN(ctx)
    HIPX(hipSetDevice(ctx->device));
    const int64_t n = a->n(), nnz = a->nnz();
    if (indptr) d2h(ctx, indptr, a->pat->rowptr.p, (n + 1) * sizeof(int32_t));
    if (indices) d2h(ctx, indices, a->pat->cols.p, nnz * sizeof(int32_t));
    if (data) d2h(ctx, data, a->vals.p, nnz * (a->dtype == LH_F64 ? 8 : 4));
    LH_API_END(ctx)
}

int lh_csr_create(lh_ctx* ctx, int64_t n, int64_t nnz, const int32_t* indptr, const int32_t* indices,
                  const void* data, int dtype, lh_csr** out) {
    if (!ctx || !out) return LH_ERR_ARG;
    *out = nullptr;
    LH_API_BEGIN(ctx)
    HIPX(hipSetDevice(ctx->device));
    if (n <= 0 || nnz < 0 || nnz > INT32_MAX) LH_THROW(LH_ERR_ARG, "bad CSR size n=%lld nnz=%lld", (long long)n, (long long)nnz);
    if (dtype != LH_F64 && dtype != LH_F32) LH_THROW(LH_ERR_ARG, "bad dtype %d", dtype);
    // validate on host: monotone indptr, columns in range, sorted rows
    if (indptr[0] != 0 || indptr[n] != nnz) LH_THROW(LH_ERR_ARG, "indptr must start at 0 and end at nnz");
    bool diag_only = true;
    for (int64_t i = 0; i < n; ++i) {
        if (indptr[i + 1] < indptr[i]) LH_THROW(LH_ERR_ARG, "indptr not monotone at row %lld", (long long)i);
        for (int32_t s = indptr[i]; s < indptr[i + 1]; ++s) {
            if (indices[s] < 0 || indices[s] >= n) LH_THROW(LH_ERR_ARG, "column index out of range at %d", s);
            if (s > indptr[i] && indices[s] <= indices[s - 1]) LH_THROW(LH_ERR_ARG, "row %lld not sorted/unique", (long long)i);
            if (indices[s] != i) diag_only = false;
        }
    }
    auto a = std::make_unique<lh_csr>();
    a->ctx = ctx;
    a->dtype = dtype;
    a->pat = std::make_shared<Pattern>();
    a->pat->n = n;
    a->pat->nnz = nnz;
    a->pat->rowptr.alloc((n + 1) * 4);
    a->pat->cols.alloc(nnz * 4);
    a->pat->diagpos.alloc(n * 4);
    a->vals.alloc(nnz * (dtype == LH_F64 ? 8 : 4));
    h2d(ctx, a->pat->rowptr.p, indptr, (n + 1) * 4);
    h2d(ctx, a->pat->cols.p, indices, nnz * 4);
    h2d(ctx, a->vals.p, data, nnz * (dtype == LH_F64 ? 8 : 4));
    k_diagpos<<<nblocks(n, 256), 256, 0, ctx->stream>>>(a->pat->rowptr.as<int32_t>(), a->pat->cols.as<int32_t>(),
                                                       a->pat->diagpos.as<int32_t>(), n);
    LH_CHECK_LAUNCH();
    a->is_diag = diag_only ? 1 : 0;
    refresh_dense_diag(ctx, a.get());
    HIPX(hipStreamSynchronize(ctx->stream));
    *out = a.release();
    LH_API_END(ctx)
}

void lh_csr_free(lh_csr* a) {
    if (!a) return;
    (void)hipSetDevice(a->ctx->device);
    (void)hipStreamSynchronize(a->ctx->stream);
    delete a;
}

int lh_csr_matmul(lh_ctx* ctx, lh_csr* a, const double* x, int nrhs, double* y) {
    if (!ctx || !a || nrhs <= 0) return LH_ERR_ARG;
    LH_API_BEGIN(ctx)
    HIPX(hipSetDevice(ctx->device));
    const int64_t n = a->n();
    DevBuf dx(n * nrhs * 8), dy(n * nrhs * 8);
    h2d(ctx, dx.p, x, n * nrhs * 8);
    if (a->dtype == LH_F64) spmm<double, double>(ctx, a, dx.as<double>(), dy.as<double>(), nrhs, nrhs);
    else spmm<float, double>(ctx, a, dx.as<double>(), dy.as<double>(), nrhs, nrhs);
    d2h(ctx, y, dy.p, n * nrhs * 8);
    LH_API_END(ctx)
}

}  // extern "C"
