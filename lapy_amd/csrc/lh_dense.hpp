// Declarations of the dense block kernels (lh_dense.hip).
#pragma once
#include "lh_internal.hpp"

namespace lh {

// ---- short-latency reciprocal and square root --------------------------------
// The hardware reciprocal (square root) estimate refined by Newton
// (Goldschmidt) steps to within an ulp or two: about half the dependent
// instructions of the correctly rounded division and square root.  For
// normal, finite positive arguments away from the exponent range's ends.
__device__ __forceinline__ double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
__device__ __forceinline__ double sqrt_nr(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double s = x * y, h = 0.5 * y;  // s ~ sqrt(x), h ~ 1 / (2 sqrt(x))
    double r = fma(-s, h, 0.5);
    s = fma(s, r, s);
    h = fma(h, r, h);
    r = fma(-s, h, 0.5);
    s = fma(s, r, s);
    h = fma(h, r, h);
    return fma(fma(-s, s, x), h, s);
}

// ---- cross-lane sums in VALU (DPP and gfx950's permlane swaps; no LDS) ----
// Every lane of the group receives the sum, with the same bits in every lane
// (each stage adds a lane's value and its partner's, and the partner adds the
// same two numbers).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
// over the 16 lanes of a DPP row
__device__ __forceinline__ double row_sum16(double v) {
    v += dpp_f64<0xB1>(v);   // quad_perm [1,0,3,2]: lane ^ 1
    v += dpp_f64<0x4E>(v);   // quad_perm [2,3,0,1]: lane ^ 2
    v += dpp_f64<0x141>(v);  // row_half_mirror: 7 - lane within 8
    v += dpp_f64<0x140>(v);  // row_mirror: 15 - lane within 16
    return v;
}
// over the 32 lanes of a half-wave
__device__ __forceinline__ double half_sum32(double v) {
    v = row_sum16(v);
    const auto lo = __builtin_amdgcn_permlane16_swap(__double2loint(v), __double2loint(v), false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap(__double2hiint(v), __double2hiint(v), false, false);
    return __hiloint2double(hi[0], lo[0]) + __hiloint2double(hi[1], lo[1]);
}
// over the 64 lanes of a wave
__device__ __forceinline__ double wave_sum64(double v) {
    v = half_sum32(v);
    const auto lo = __builtin_amdgcn_permlane32_swap(__double2loint(v), __double2loint(v), false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap(__double2hiint(v), __double2hiint(v), false, false);
    return __hiloint2double(hi[0], lo[0]) + __hiloint2double(hi[1], lo[1]);
}

// C (p x q, ldc, FP64) = alpha * U^T V over n rows; sym (p == q): the
// product is known to be symmetric, only its upper 64 x 64 tiles are formed
// and the lower ones mirror them
template <typename T>
void gram(lh_ctx* ctx, int64_t n, const T* U, int ldu, int p, const T* V, int ldv, int q, double* C, int ldc,
          DevBuf& ws, double alpha = 1.0, bool sym = false);
// Y = alpha * U C + beta * Y
template <typename T>
void gemm_nn(lh_ctx* ctx, int64_t n, int p, int q, const T* U, int ldu, const double* C, int ldc, T* Y, int ldy,
             double alpha, double beta);
// matrix-core forms of the two (lh_mfma.hip; T = double or float)
template <typename T>
void gram_mfma(lh_ctx* ctx, int64_t n, const T* U, int ldu, int p, const T* V, int ldv, int q, double* C, int ldc,
               DevBuf& ws, double alpha, bool sym);
template <typename T>
void gemm_mfma(lh_ctx* ctx, int64_t n, int p, int q, const T* U, int ldu, const double* C, int ldc, T* Y, int ldy,
               double alpha, double beta, bool upper = false);
// X (n x p, p <= 256) <- alpha X C in place; upper: C upper triangular
template <typename T>
void gemm_mfma_inplace(lh_ctx* ctx, int64_t n, int p, T* X, int ldx, const double* C, int ldc, double alpha,
                       bool upper);
size_t rr_work_doubles(int N, int nev);
// standard symmetric Rayleigh-Ritz (orthonormal basis), LDS-resident, N <= kRRStdMaxN
constexpr int kRRStdMaxN = 192;
size_t rr_std_work_doubles(int N, int nev);
void rr_std(lh_ctx* ctx, int N, int nev, const double* G, int ldg, double* C, double* theta, double* work,
            int* status, unsigned long long* prof = nullptr);
void rr_solve(lh_ctx* ctx, int N, int nev, double* GA, double* GM, double* C, double* theta, double* work,
              int* status);
// the same generalized problem through rocSOLVER (sygvd) for trial bases
// wider than the LDS solvers take; work: rr_wide_work_doubles(N) doubles
size_t rr_wide_work_doubles(int N);
void rr_solve_wide(lh_ctx* ctx, int N, int nev, const double* GA, const double* GM, double* C, double* theta,
                   double* work, int* status);
// widest Gram cholinv factorises (one workgroup, the factor in LDS)
constexpr int kCholMax = 128;
void cholinv(lh_ctx* ctx, int m, double* G, double* T, double shift, int* status);
template <typename T>
void residual(lh_ctx* ctx, int64_t n, int m, const T* AX, const T* MX, int ldx, const double* theta, T* R, int ldr);
template <typename T>
void colsq(lh_ctx* ctx, int64_t n, int m, const T* X, int ld, double* out, DevBuf& ws);
// residual() fused with colsq() of R (into r2) and of MX (into mx2)
template <typename T>
void residual_norms(lh_ctx* ctx, int64_t n, int m, const T* AX, const T* MX, int ldx, const double* theta, T* R,
                    int ldr, double* r2, double* mx2, DevBuf& ws);
template <typename T>
void colscale(lh_ctx* ctx, int64_t n, int m, T* X, int ld, const double* s);
template <typename T>
void rowscale(lh_ctx* ctx, int64_t n, int m, T* X, int ld, const double* dinv);
// dst[:, c] = src[:, perm[c]], c < m; perm: m int32 on the device
template <typename T>
void gather_cols(lh_ctx* ctx, int64_t n, int m, const T* src, int lds, T* dst, int ldd, const int32_t* perm);
template <typename S, typename D>
void copy_block(lh_ctx* ctx, int64_t n, int m, const S* src, int lds, D* dst, int ldd);
template <typename T>
void randn(lh_ctx* ctx, int64_t n, int m, T* X, int ld, uint64_t seed);
// dinv = 1 / (alpha * diag(B) + beta * diag(A))
void jacobi_diag(lh_ctx* ctx, const lh_csr* A, const lh_csr* B, double alpha, double beta, double* dinv);

// values of K = beta A + alpha B on A's pattern (B diagonal or on that pattern)
void form_shifted(lh_ctx* ctx, const lh_csr* A, const lh_csr* B, double alpha, double beta, DevBuf& kv);

// Chebyshev-accelerated Jacobi preconditioner for K = A - sigma B (lh_poly.hip)
struct ChebPrecond {
    lh_ctx* ctx = nullptr;
    std::shared_ptr<Pattern> pat;
    int64_t n = 0;
    int degree = 1;
    int mmax = 0;
    bool half = false;  // FP16 storage of the work blocks (FP32 arithmetic)
    double bmax = 2.0, amin = 0.03;
    DevBuf kv, dinv, kv32, dinv32, R, D0, D1, sc;
    // ELL copy of the FP32 matrix (FP16 steps): ell_w slots per row, 0: none
    static constexpr int kEllMaxW = 32;
    int ell_w = 0;
    DevBuf ecols, ekv;
    void setup(lh_ctx* c, const lh_csr* A, const lh_csr* B, double sigma, int deg, double ratio, int64_t n_,
               int mmax_, bool half_storage);
    // W (n x m, leading dimension ldw) <- p(D^-1 K) D^-1 W; colnorm2 (device,
    // m doubles) holds the squared column norms of W (needed for FP16 storage)
    template <typename WT>  // double or float blocks
    void apply(WT* W, int ldw, int m, const double* colnorm2);
    // average duration (us) of one Chebyshev step on an n x m block
    double time_step(int m, int iters);

   private:
    template <typename WT>
    void init(const WT* W, int ldw, int m, const double* colnorm2, void** xk, void** xp);
    void step(int m, const void* bvec, const void* xk, void* xp, double omega);
};

// learned message-passing preconditioner (lh_mpnn.hip) for K = beta A + alpha B
struct MpnnRun;
struct LearnedPrecond {
    lh_mpnn* net = nullptr;
    std::unique_ptr<MpnnRun> run;
    LearnedPrecond();
    ~LearnedPrecond();
    void setup(lh_ctx* c, lh_mpnn* net, const lh_csr* A, const lh_csr* B, double alpha, double beta);
    // W (n x m, leading dimension ldw) <- P W
    void apply(double* W, int ldw, int m);
    void apply(float* W, int ldw, int m);
};

// locality reordering of a symmetric pattern: sorted Morton codes of a mesh's
// vertices, or breadth-first levels (lh_reorder.hip)
struct Reorder {
    DevBuf perm;   // new row -> old row
    DevBuf iperm;  // old row -> new row
    DevBuf src;    // new nonzero slot -> old slot
    std::shared_ptr<Pattern> pat;  // permuted pattern
    int nlevels = 0;
    void build(lh_ctx* ctx, const Pattern& p);
    // P a P^T on the permuted pattern (diagonal matrices: permuted diagonal)
    std::unique_ptr<lh_csr> apply(lh_ctx* ctx, const lh_csr* a, std::shared_ptr<Pattern> for_pattern = nullptr,
                                  const Reorder* diag_reorder = nullptr) const;
    // out[i, :] = in[iperm[i], :]
    void unpermute(lh_ctx* ctx, const double* in, int ldi, double* out, int ldo, int64_t n, int m) const;
};

// the locality order of a pattern: built on the first request (outside any
// call arena) and kept with the pattern; concurrent first requests build it
// once
std::shared_ptr<Reorder> pattern_reorder(lh_ctx* ctx, Pattern& p);

}  // namespace lh
