// Internal declarations shared by the HIP sources of liblapyhip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/lapyhip.h"

// --------------------------------------------------------------------------
// errors
// --------------------------------------------------------------------------
struct lh_error {
    int code;
    std::string msg;
};

#define LH_THROW(code, ...)                                                   \
    do {                                                                      \
        char _b[512];                                                         \
        snprintf(_b, sizeof(_b), __VA_ARGS__);                                \
        throw lh_error{(code), std::string(_b)};                              \
    } while (0)

#define HIPX(expr)                                                            \
    do {                                                                      \
        hipError_t _e = (expr);                                               \
        if (_e != hipSuccess) {                                               \
            int _c = (_e == hipErrorOutOfMemory) ? LH_ERR_OOM : LH_ERR_HIP;   \
            LH_THROW(_c, "%s failed: %s (%s:%d)", #expr,                      \
                     hipGetErrorString(_e), __FILE__, __LINE__);              \
        }                                                                     \
    } while (0)

#define LH_CHECK_LAUNCH() HIPX(hipGetLastError())

// --------------------------------------------------------------------------
// device buffers (RAII)
// --------------------------------------------------------------------------
// A bump allocator over one device allocation.  A thread that sets
// lh_thread_arena takes its solver buffers from it (no hipMalloc/hipFree,
// which synchronise the device, between the problems of a batch) and resets
// `top` to a mark once every buffer of a problem is gone.
struct Arena {
    char* base = nullptr;
    size_t cap = 0, top = 0;
    void* take(size_t n) {
        n = (n + 255) & ~size_t(255);
        if (!base || top + n > cap) return nullptr;
        void* q = base + top;
        top += n;
        return q;
    }
};
extern thread_local Arena* lh_thread_arena;

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    bool pooled = false;      // taken from the thread's arena (not freed here)
    bool persistent = false;  // never taken from an arena (outlives a problem)
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes), pooled(o.pooled), persistent(o.persistent) {
        o.p = nullptr;
        o.bytes = 0;
        o.pooled = false;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p; bytes = o.bytes; pooled = o.pooled;
            o.p = nullptr; o.bytes = 0; o.pooled = false;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t n) {
        release();
        if (n == 0) n = 8;
        if (!persistent && lh_thread_arena) {
            if (void* q = lh_thread_arena->take(n)) {
                p = q;
                bytes = n;
                pooled = true;
                return;
            }
        }
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            p = nullptr;
            LH_THROW(LH_ERR_OOM, "hipMalloc(%zu bytes) failed: %s", n, hipGetErrorString(e));
        }
        bytes = n;
    }
    // grow-only reallocation (contents not preserved)
    void ensure(size_t n) { if (n > bytes) alloc(n); }
    void release() {
        if (p && !pooled) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        pooled = false;
    }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

// allocations in this scope bypass the thread's arena (results that outlive
// the problem)
struct ArenaOff {
    Arena* saved;
    ArenaOff() : saved(lh_thread_arena) { lh_thread_arena = nullptr; }
    ~ArenaOff() { lh_thread_arena = saved; }
};

// --------------------------------------------------------------------------
// handles
// --------------------------------------------------------------------------
// tuned workgroup size of the ELL SpMM for one problem shape (lh_eigs_autotune,
// lh_tune_set), looked up by the matrix's (n, nnz); 256 threads when absent.
// k and ncol (the solver's block width) record what it was tuned on.
// Shared by a context and its batch workers; not changed while solves run.
struct SpmmTune {
    int64_t n, nnz;
    int k, ncol, block;
};
struct TuneTable {
    std::mutex mu;
    std::vector<SpmmTune> spmm;
};

struct lh_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;
    int n_cu = 256;
    DevBuf scratch;      // general scratch (grow-only)
    DevBuf small;        // small reduction scratch
    DevBuf pinned_dummy;
    // batched eigensolver: worker contexts (own streams) and their arenas
    std::vector<lh_ctx*> workers;
    DevBuf arena_mem;
    Arena arena;
    int tuned_workers = 0;  // 0: not tuned
    lh_mpnn* mpnn = nullptr;  // network used by LH_PRECOND_LEARNED (not owned)
    void* host_small = nullptr;   // pinned host memory (hipHostMalloc): the PCG's status polling, the
                                  // eigensolver's soft-locking uploads (never both at once on one context)
    size_t host_small_bytes = 0;
    void* host_stage = nullptr;   // pinned staging of d2h / h2d / d2h_n
    size_t host_stage_bytes = 0;
    void* blas = nullptr;  // rocBLAS handle of the wide Rayleigh-Ritz solver, made on first use
    int asm_order = 0;     // LH_ASM_ORDER_*: row order of the triangle assembly
    std::shared_ptr<TuneTable> tune = std::make_shared<TuneTable>();
};

// frees ctx->blas (lh_rrlib.hip)
void release_blas(lh_ctx* ctx);

// incidence lists, CSR pattern and AoS coordinates of a mesh, built once
// (lh_assembly.hip)
void lh_build_topology(lh_ctx* ctx, lh_mesh* m);

// device address of lh_block_gram_noalloc's fixed scratch on `device`, looked
// up once per device with that device current (lh_mfma.hip; called by
// lh_ctx_create)
void noalloc_scratch_init(int device);

namespace lh {
struct Reorder;
}
// locality-ordered row schedule of the triangle assembly (lh_assembly.hip)
struct AsmSched;
// CSR pattern (rows sorted, int32), shared by A and consistent B.
struct Pattern {
    int64_t n = 0, nnz = 0;
    DevBuf rowptr;   // n+1 int32
    DevBuf cols;     // nnz int32
    DevBuf diagpos;  // n int32 (-1: row without diagonal)
    // locality reordering of this pattern, built by the first solve (or
    // assembly) that wants it and kept for later ones on the same pattern
    // (lh::pattern_reorder)
    std::shared_ptr<lh::Reorder> ro;
    std::mutex ro_mu;
    // optional per-row sort key of the locality order (the Morton code of a
    // mesh vertex, set with the mesh's pattern); without it the order is a
    // breadth-first search of the pattern
    DevBuf order_key;
};

struct lh_csr {
    lh_ctx* ctx = nullptr;
    std::shared_ptr<Pattern> pat;
    int dtype = LH_F64;
    int is_diag = 0;     // pattern holds only diagonal entries
    DevBuf vals;         // nnz values
    DevBuf dense_diag;   // for is_diag: n values (0 where the pattern has no entry)
    int64_t n() const { return pat->n; }
    int64_t nnz() const { return pat->nnz; }
};

struct lh_mesh {
    lh_ctx* ctx = nullptr;
    int64_t nv = 0, ne = 0;
    int nodes = 3;
    int vdtype = LH_F64;
    DevBuf x, y, z;      // SoA coordinates
    DevBuf t;            // int32 ne x nodes
    // topology (built lazily)
    int64_t n = -1;                      // max(t)+1
    DevBuf inc_ptr, inc;                 // vertex -> (e<<2 | local), ascending
    DevBuf incv;                         // per incidence entry: int4 of the element's
                                         // vertices (triangles: 3 + local index)
    DevBuf xyz;                          // AoS copy (x, y, z, 0) in the vertex dtype
    std::shared_ptr<Pattern> full;       // consistent pattern
    std::shared_ptr<Pattern> lumped;     // diagonal-of-used-vertices pattern
    std::shared_ptr<AsmSched> sched;     // built on first use by a large triangle mesh
    int sched_state = 0;                 // 0: not tried, 1: built, -1: mesh not eligible
    int64_t numeric_runs = 0;            // numeric assembly launches so far
};

struct lh_eig {
    lh_ctx* ctx = nullptr;
    int64_t n = 0;
    int k = 0;
    std::vector<double> evals;
    DevBuf evecs;        // n x k row-major on device
};

// --------------------------------------------------------------------------
// shared helpers (defined in lh_core.hip)
// --------------------------------------------------------------------------
namespace lh {

inline unsigned nblocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// MI355X places workgroup b of a launch on XCD (b mod 8), and each of the 8
// XCDs has its own L2.  Row-blocked gather kernels launch a multiple of 8
// workgroups and renumber them so that each XCD walks one contiguous eighth
// of the rows: neighbouring rows (which gather the same block rows) then
// share an L2.
constexpr int kXCDs = 8;
inline unsigned nblocks_xcd(int64_t n, int bs) { return ((nblocks(n, bs) + kXCDs - 1) / kXCDs) * kXCDs; }
__device__ __forceinline__ int64_t xcd_block() {
    const unsigned per = gridDim.x / kXCDs;
    return (int64_t)(blockIdx.x % kXCDs) * per + blockIdx.x / kXCDs;
}

// exclusive scan of n int32 values in place into out (n+1 entries, out[n]=total)
void exclusive_scan_i32(lh_ctx* ctx, const int32_t* in, int32_t* out, int64_t n);
// host-synchronous small copy helpers
void d2h(lh_ctx* ctx, void* dst, const void* src, size_t bytes);
// several device -> host copies behind one stream synchronisation
struct Xfer {
    void* host;
    const void* dev;
    size_t bytes;
};
void d2h_n(lh_ctx* ctx, const Xfer* x, int cnt);
void h2d(lh_ctx* ctx, void* dst, const void* src, size_t bytes);
void memset_d(lh_ctx* ctx, void* dst, int v, size_t bytes);

std::shared_ptr<Pattern> make_diag_pattern(lh_ctx* ctx, const Pattern& full, int64_t n,
                                           const int32_t* used_flag);

// SpMM / SpMV: Y = a * X  (X, Y row-major n x ncol, leading dimension ld)
template <typename VT, typename XT>
void spmm(lh_ctx* ctx, const lh_csr* a, const XT* X, XT* Y, int ncol, int ld);
template <typename VT, typename XT>
void spmm_ld(lh_ctx* ctx, const lh_csr* a, const XT* X, XT* Y, int ncol, int ldx, int ldy);
// YA = a X, YB = b X (one pass when a and b share a pattern)
template <typename VT, typename XT>
void spmm_pair(lh_ctx* ctx, const lh_csr* a, const lh_csr* b, const XT* X, XT* YA, XT* YB, int ncol, int ldx,
               int ldy);

void refresh_dense_diag(lh_ctx* ctx, lh_csr* a);
// FP32 copy of an FP64 matrix on the same pattern
std::unique_ptr<lh_csr> csr_to_f32(lh_ctx* ctx, const lh_csr* s);

// ELL copy of one matrix or of two on one pattern, for the block products of
// the eigensolver: W slots per row (W % 8 == 0, row-major), a padding slot
// points at the row itself with value 0.  Without row pointers in the way,
// the slots' columns and values and eight gathers go out together.
struct Ell {
    static constexpr int kMaxW = 32;
    int64_t n = 0, nnz = 0;
    int W = 0;          // 0: not built (rows too long, or a diagonal matrix)
    bool two = false;   // vb holds the second matrix
    int dtype = LH_F64;
    DevBuf cols, va, vb;
    void build(lh_ctx* ctx, const lh_csr* a, const lh_csr* b);
};
// YA = a X (and YB = b X when the ELL holds two matrices and YB is given);
// which = 1: YB = b X.  Returns false (nothing launched) when the copy is not
// built or the block is not 16-byte vectorisable; the caller then uses CSR.
// block: threads per workgroup (64, 128 or 256); 0 takes the context's tuned
// size for this shape (256 when untuned).
template <typename VT, typename XT>
bool spmm_ell(lh_ctx* ctx, const Ell& e, const XT* X, XT* YA, XT* YB, int ncol, int ldx, int ldy, int which = 0,
              int block = 0);
// the tuned ELL SpMM workgroup size for a matrix shape (0: none)
int tuned_spmm_block(const lh_ctx* ctx, int64_t n, int64_t nnz);

// 16-byte vector of a block row: 4 floats or 2 doubles per lane
template <typename T> struct Vec16;
template <> struct Vec16<float> {
    using type = float4;
    static constexpr int n = 4;
};
template <> struct Vec16<double> {
    using type = double2;
    static constexpr int n = 2;
};
template <typename T>
__device__ __forceinline__ void vfma(typename Vec16<T>::type& acc, T w, const typename Vec16<T>::type& x);
template <>
__device__ __forceinline__ void vfma<float>(float4& acc, float w, const float4& x) {
    acc.x += w * x.x; acc.y += w * x.y; acc.z += w * x.z; acc.w += w * x.w;
}
template <>
__device__ __forceinline__ void vfma<double>(double2& acc, double w, const double2& x) {
    acc.x += w * x.x; acc.y += w * x.y;
}
template <typename T>
__device__ __forceinline__ typename Vec16<T>::type vzero() {
    typename Vec16<T>::type z;
    z.x = T(0); z.y = T(0);
    if constexpr (Vec16<T>::n == 4) { z.z = T(0); z.w = T(0); }
    return z;
}
template <typename T>
__device__ __forceinline__ void vadd(typename Vec16<T>::type& a, const typename Vec16<T>::type& b) {
    a.x += b.x; a.y += b.y;
    if constexpr (Vec16<T>::n == 4) { a.z += b.z; a.w += b.w; }
}
// acc = sum_s vals[s] * X[cols[s], c0 .. c0+VEC) over one CSR row, two
// independent accumulators so that two gathers are in flight per lane
template <typename VT, typename T>
__device__ __forceinline__ typename Vec16<T>::type row_gather(const int32_t* __restrict__ cols,
                                                              const VT* __restrict__ vals, int32_t r0, int32_t r1,
                                                              const T* __restrict__ X, int ldx, int c0) {
    using V = typename Vec16<T>::type;
    V a0 = vzero<T>(), a1 = vzero<T>();
    int32_t s = r0;
    for (; s + 1 < r1; s += 2) {
        const int32_t j0 = cols[s], j1 = cols[s + 1];
        const T w0 = (T)vals[s], w1 = (T)vals[s + 1];
        const V x0 = *reinterpret_cast<const V*>(X + (int64_t)j0 * ldx + c0);
        const V x1 = *reinterpret_cast<const V*>(X + (int64_t)j1 * ldx + c0);
        vfma<T>(a0, w0, x0);
        vfma<T>(a1, w1, x1);
    }
    if (s < r1) {
        const V x0 = *reinterpret_cast<const V*>(X + (int64_t)cols[s] * ldx + c0);
        vfma<T>(a0, (T)vals[s], x0);
    }
    vadd<T>(a0, a1);
    return a0;
}
// lanes per row for a vectorised block row of ncol values (power of two)
inline int vec_lanes(int ncol, int vec) {
    int need = (ncol + vec - 1) / vec, l = 1;
    while (l < need) l <<= 1;
    return l;
}

}  // namespace lh

// C-ABI guard
#define LH_API_BEGIN(ctx) try {
#define LH_API_END(ctx)                                                       \
    }                                                                         \
    catch (const lh_error& e) {                                               \
        if (ctx) (ctx)->last_error = e.msg;                                   \
        return e.code;                                                        \
    }                                                                         \
    catch (const std::exception& e) {                                         \
        if (ctx) (ctx)->last_error = e.what();                                \
        return LH_ERR_HIP;                                                    \
    }                                                                         \
    return LH_OK;
