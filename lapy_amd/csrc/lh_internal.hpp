// Internal declarations shared by the HIP sources of liblapyhip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <memory>
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
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t n) {
        release();
        if (n == 0) n = 8;
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
        if (p) { (void)hipFree(p); p = nullptr; bytes = 0; }
    }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

// --------------------------------------------------------------------------
// handles
// --------------------------------------------------------------------------
struct lh_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;
    int n_cu = 256;
    DevBuf scratch;      // general scratch (grow-only)
    DevBuf small;        // small reduction scratch
    DevBuf pinned_dummy;
    void* host_small = nullptr;   // pinned host staging (hipHostMalloc)
    size_t host_small_bytes = 0;
};

// CSR pattern (rows sorted, int32), shared by A and consistent B.
struct Pattern {
    int64_t n = 0, nnz = 0;
    DevBuf rowptr;   // n+1 int32
    DevBuf cols;     // nnz int32
    DevBuf diagpos;  // n int32 (-1: row without diagonal)
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
    std::shared_ptr<Pattern> full;       // consistent pattern
    std::shared_ptr<Pattern> lumped;     // diagonal-of-used-vertices pattern
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

// exclusive scan of n int32 values in place into out (n+1 entries, out[n]=total)
void exclusive_scan_i32(lh_ctx* ctx, const int32_t* in, int32_t* out, int64_t n);
// host-synchronous small copy helpers
void d2h(lh_ctx* ctx, void* dst, const void* src, size_t bytes);
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
