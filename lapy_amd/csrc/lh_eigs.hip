// Generalized symmetric eigensolver  A x = lambda B x  (smallest k pairs).
//
// Replaces Solver.eigs (lapy/solver.py:667-716), which factorises A - sigma*B
// with SuperLU and runs ARPACK's shift-invert Lanczos (eigsh).  Here: block
// LOBPCG with soft guard vectors, on the GPU, FP64:
//
//   S = [X | W | P]  (n x 3m, row-major, one array so that a single Gram and a
//                    single GEMM cover the whole trial subspace)
//   AS, MS           A*S and B*S, recomputed with SpMM after every change of a
//                    block (more accurate than the usual GEMM recurrences and
//                    not slower: the SpMM reads the block once, a GEMM update
//                    reads the 3m-wide S)
//   W = D^-1 (AX - BX Theta),  D = diag(A - sigma*B)      (Jacobi)
//   W <- W - X (BX)^T W (twice),  Cholesky-QR in the B inner product
//   P <- P - X (BX)^T P - W (BW)^T P,  Cholesky-QR (dropped on breakdown)
//   Rayleigh-Ritz on span(S) with the one-workgroup FP64 solver (lh_rr.hip)
//   X <- S C,  P <- S[:, m:] C[m:]
//
// The host waits for the device at the residual norms (which also bring the
// previous Rayleigh-Ritz's status and Ritz values) and at each Cholesky-QR
// pass's pivot check; with the Jacobi and Chebyshev preconditioners
// everything else is queued without waiting.
//
// Convergence: ||A x_j - theta_j B x_j||_2 / ||B x_j||_2 <= tol * scale_j with
// scale_j = max(|theta_j|, 1e-2 * |theta_{k-1}|) for the k wanted columns.
#include "lh_dense.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <thread>

using namespace lh;

namespace {

double now_us() {
    return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// the polynomial skips the wanted pairs whose relative residual is at most
// this many times the tolerance (they keep a plain Jacobi direction)
constexpr double kPolyLockFactor = 1.0;

// T: the block type.  double for the FP64 solve; float for the first phase
// of the mixed-precision solve (FP32 blocks, FP32 SpMM, FP32 matrix-core
// Grams and updates, FP64 Rayleigh-Ritz and Cholesky factors), which stops
// at a looser tolerance and hands its X block to an FP64 solve.
template <typename T>
struct Lobpcg {
    lh_ctx* ctx;
    const lh_csr* A;  // FP64 matrices (reordered when the solver reorders)
    const lh_csr* B;
    const lh_csr* Am = nullptr;  // the matrices of the block products: A, B or their FP32 copies
    const lh_csr* Bm = nullptr;
    std::unique_ptr<lh_csr> A32, B32;
    int64_t n;
    int k, m, ld;
    double sigma, tol;
    int max_iter, verbose;
    int poly_degree = 0;
    double poly_ratio = 0.0;
    bool poly_half = false;
    int precond = LH_PRECOND_JACOBI;
    bool no_fallback = false;
    // first phase of a mixed-precision solve: also stop when the worst
    // residual has stopped falling (FP32 rounding floor)
    bool phase_one = false;
    const T* x0 = nullptr;  // start block of a second phase (n x m, leading dimension x0_ld), else random
    int x0_ld = 0;
    LearnedPrecond learned;
    ChebPrecond own_cheb;
    ChebPrecond* cheb = &own_cheb;  // a second phase shares the first one's
    bool cheb_ready = false;
    // locality reordering (solver-internal)
    std::shared_ptr<Reorder> ro;      // the pattern's (kept with it)
    const Reorder* ext_ro = nullptr;  // a second phase's: the first phase's order
    std::unique_ptr<lh_csr> Ap, Bp;
    bool reordered = false;
    uint64_t seed;
    DevBuf S, S2, AS, MS, dinv;
    DevBuf GA, GM, Cm, Gs, Ts, theta_d, rrwork, status, cs1, cs2, wsg, wsc;
    DevBuf lockbuf;  // soft locking of the polynomial: column norms (m doubles), then the column order (m int32)
    std::vector<double> theta, theta_new, res;
    lh_stats st{};

    T* Sp() { return S.as<T>(); }
    T* ASp() { return AS.as<T>(); }
    T* MSp() { return MS.as<T>(); }

    Ell ell;  // ELL copy of Am (and Bm when it shares Am's pattern)

    void spmm_AB(const T* X, T* AX, T* MX, int ncol) {
        if (spmm_ell<T, T>(ctx, ell, X, AX, ell.two ? MX : nullptr, ncol, ld, ld)) {
            if (!ell.two) spmm_ld<T, T>(ctx, Bm, X, MX, ncol, ld, ld);
        } else {
            spmm_pair<T, T>(ctx, Am, Bm, X, AX, MX, ncol, ld, ld);
        }
        st.spmm_calls++;
    }
    void spmm_B(const T* X, T* MX, int ncol) {
        if (!spmm_ell<T, T>(ctx, ell, X, nullptr, MX, ncol, ld, ld, 1)) spmm_ld<T, T>(ctx, Bm, X, MX, ncol, ld, ld);
        st.spmm_calls++;
    }
    int read_status() {
        int s = 0;
        d2h(ctx, &s, status.p, sizeof(int));
        return s;
    }

    // X block (cols [off, off+c)) <- X T with T^T (X^T B X) T = I; B X must be
    // current on entry.  A second pass follows when the first one met a small
    // pivot (Cholesky-QR2).  Returns false when the block is numerically rank
    // deficient.  A X and B X of the block are recomputed afterwards.
    bool cholqr(int off, int c, bool recompute_A) {
        if (c > kCholMax) {
            // wider than one factorisation: the first half (W of [W P]),
            // then the second half projected against it (twice) and
            // factorised on its own, each half split again while wider
            // than kCholMax.  A rank-deficient second half fails as the
            // whole block would.
            const int c1 = (c + 1) / 2, c2 = c - c1;
            if (!cholqr(off, c1, recompute_A)) return false;
            project_out(off, c1, off + c1, c2);
            project_out(off, c1, off + c1, c2);
            spmm_B(Sp() + off + c1, MSp() + off + c1, c2);
            return cholqr(off + c1, c2, recompute_A);
        }
        T* X = Sp() + off;
        T* MX = MSp() + off;
        for (int pass = 0; pass < 2; ++pass) {
            gram<T>(ctx, n, X, ld, c, MX, ld, c, Gs.as<double>(), c, wsg, 1.0, true);
            cholinv(ctx, c, Gs.as<double>(), Ts.as<double>(), 0.0, status.as<int>());
            double stv[2];
            d2h(ctx, stv, status.p, sizeof(stv));
            if (reinterpret_cast<int*>(stv)[0]) return false;
            gemm_mfma_inplace(ctx, n, c, X, ld, Ts.as<double>(), c, 1.0, true);  // T is upper triangular
            if (stv[1] > 1e-2) break;
            gemm_mfma_inplace(ctx, n, c, MX, ld, Ts.as<double>(), c, 1.0, true);
        }
        if (recompute_A) spmm_AB(X, ASp() + off, MX, c);
        else spmm_B(X, MX, c);
        return true;
    }

    // Y -= X C with C = (B X)^T Y  (X block at xoff, Y block at yoff, widths cx, cy)
    void project_out(int xoff, int cx, int yoff, int cy) {
        gram<T>(ctx, n, MSp() + xoff, ld, cx, Sp() + yoff, ld, cy, Cm.as<double>(), cy, wsg);
        gemm_nn<T>(ctx, n, cx, cy, Sp() + xoff, ld, Cm.as<double>(), cy, Sp() + yoff, ld, -1.0, 1.0);
    }

    void normalize_cols(int off, int c) {
        colsq<T>(ctx, n, c, Sp() + off, ld, cs1.as<double>(), wsc);
        std::vector<double> h(c);
        d2h(ctx, h.data(), cs1.p, c * 8);
        for (int j = 0; j < c; ++j) h[j] = h[j] > 0 ? 1.0 / std::sqrt(h[j]) : 1.0;
        h2d(ctx, cs1.p, h.data(), c * 8);
        colscale<T>(ctx, n, c, Sp() + off, ld, cs1.as<double>());
    }

    // Rayleigh-Ritz on the first N columns of S; returns false on breakdown
    // The blocks of S are B-orthonormalised (X by the previous Rayleigh-Ritz,
    // W and P by projection + Cholesky-QR), so S^T B S = I to rounding and
    // the small problem is a standard one; the generalized path (Cholesky of
    // S^T B S, through rocSOLVER) is kept for trial bases wider than the LDS
    // solver takes.
    // launches only; the status and the Ritz values (theta_d) are read by
    // rayleigh_ritz or, inside the iteration, with the next residual norms
    void rr_issue(int N) {
        gram<T>(ctx, n, Sp(), ld, N, ASp(), ld, N, GA.as<double>(), N, wsg, 1.0, true);
        if (N <= kRRStdMaxN) {
            rr_std(ctx, N, m, GA.as<double>(), N, Cm.as<double>(), theta_d.as<double>(), rrwork.as<double>(),
                   status.as<int>());
        } else {
            gram<T>(ctx, n, Sp(), ld, N, MSp(), ld, N, GM.as<double>(), N, wsg, 1.0, true);
            rr_solve_wide(ctx, N, m, GA.as<double>(), GM.as<double>(), Cm.as<double>(), theta_d.as<double>(),
                          rrwork.as<double>(), status.as<int>());
        }
        st.rr_calls++;
    }
    bool rr_pending = false;  // an in-loop Rayleigh-Ritz whose status and theta are not read yet
    bool rayleigh_ritz(int N) {
        double t0 = now_us();
        rr_issue(N);
        int s = 0;
        theta_new.resize(m);
        const Xfer x[2] = {{&s, status.p, sizeof(int)}, {theta_new.data(), theta_d.p, (size_t)m * 8}};
        d2h_n(ctx, x, 2);
        st.t_rr_us += now_us() - t0;
        if (s) return false;  // theta keeps the last good Ritz values
        theta.swap(theta_new);
        return true;
    }
    // reads a pending in-loop Rayleigh-Ritz's status and theta together with
    // the residual norms (one synchronisation)
    void read_residuals(std::vector<double>& r2, std::vector<double>& mx2, int it) {
        const Xfer rx[2] = {{r2.data(), cs1.p, (size_t)m * 8}, {mx2.data(), cs2.p, (size_t)m * 8}};
        if (!rr_pending) {
            d2h_n(ctx, rx, 2);
            return;
        }
        int s = 0;
        theta_new.resize(m);
        const Xfer x[4] = {rx[0], rx[1], {&s, status.p, sizeof(int)}, {theta_new.data(), theta_d.p, (size_t)m * 8}};
        d2h_n(ctx, x, 4);
        rr_pending = false;
        if (s) LH_THROW(LH_ERR_NUMERIC, "Rayleigh-Ritz failed at iteration %d", it - 1);
        theta.swap(theta_new);
    }

    // out_evals / out_vecs: the result (a second phase or an FP64 solve);
    // a first phase leaves its X block in S
    void run(double* out_evals, DevBuf* out_vecs) {
        double t_start = now_us();
        if (!ext_ro && n >= 4096 && (B->is_diag || B->pat == A->pat)) {
            ro = pattern_reorder(ctx, *A->pat);
            Ap = ro->apply(ctx, A);
            Bp = ro->apply(ctx, B, B->is_diag ? nullptr : ro->pat);
            A = Ap.get();
            B = Bp.get();
            reordered = true;
        }
        if constexpr (sizeof(T) == 8) {
            if (A->dtype != LH_F64 || B->dtype != LH_F64) LH_THROW(LH_ERR_ARG, "the eigensolver needs FP64 matrices");
            Am = A;
            Bm = B;
        } else {
            A32 = csr_to_f32(ctx, A);
            B32 = csr_to_f32(ctx, B);
            Am = A32.get();
            Bm = B32.get();
        }
        ell.build(ctx, Am, Bm);
        ld = 3 * m;
        S.alloc(n * ld * sizeof(T));
        S2.alloc(n * ld * sizeof(T));
        AS.alloc(n * ld * sizeof(T));
        MS.alloc(n * ld * sizeof(T));
        dinv.alloc(n * 8);
        const int N3 = 3 * m;
        GA.alloc(N3 * N3 * 8);
        GM.alloc(N3 * N3 * 8);
        Cm.alloc(N3 * std::max(N3, m) * 8);
        Gs.alloc(4 * m * m * 8);
        Ts.alloc(4 * m * m * 8);
        theta_d.alloc(N3 * 8);
        rrwork.alloc(std::max({rr_work_doubles(N3, m), rr_std_work_doubles(N3, m), rr_wide_work_doubles(N3)}) * 8);
        status.alloc(64);
        cs1.alloc(N3 * 8);
        cs2.alloc(N3 * 8);
        lockbuf.alloc(m * 12);
        theta.assign(m, 0.0);
        res.assign(m, 0.0);
        HIPX(hipMemsetAsync(S.p, 0, n * ld * sizeof(T), ctx->stream));
        jacobi_diag(ctx, A, B, -sigma, 1.0, dinv.as<double>());
        if (poly_degree > 1 && !cheb_ready) cheb->setup(ctx, A, B, sigma, poly_degree, poly_ratio, n, m, poly_half);
        bool learned_on = false;
        if (precond == LH_PRECOND_LEARNED) {
            learned.setup(ctx, ctx->mpnn, A, B, -sigma, 1.0);  // K = A - sigma B
            learned_on = true;
        }
        constexpr int kStallWindow = 16;
        double hist[kStallWindow] = {};  // residual measure of the last iterations

        // start block
        if (x0) copy_block<T, T>(ctx, n, m, x0, x0_ld, Sp(), ld);
        else randn<T>(ctx, n, m, Sp(), ld, seed);
        spmm_B(Sp(), MSp(), m);
        if (!cholqr(0, m, true)) LH_THROW(LH_ERR_NUMERIC, "start block is rank deficient");
        if (!rayleigh_ritz(m)) LH_THROW(LH_ERR_NUMERIC, "initial Rayleigh-Ritz failed");
        gemm_mfma_inplace<T>(ctx, n, m, Sp(), ld, Cm.as<double>(), m, 1.0, false);
        spmm_AB(Sp(), ASp(), MSp(), m);
        st.t_setup_us = now_us() - t_start;

        bool have_p = false;
        int it = 0;
        double t_iter0 = now_us();
        std::vector<double> r2(m), mx2(m);
        std::vector<char> conv(m, 0);
        std::vector<int32_t> perm;
        std::vector<double> gm_hist;  // first phase: residual history
        constexpr int kMinFp32Iters = 6;
        for (it = 0; it < max_iter; ++it) {
            // residual into W and the convergence test
            residual_norms<T>(ctx, n, m, ASp(), MSp(), ld, theta_d.as<double>(), Sp() + m, ld, cs1.as<double>(),
                              cs2.as<double>(), wsc);
            read_residuals(r2, mx2, it);
            const double floor = 1e-2 * std::fabs(theta[k - 1]);
            int nconv = 0;
            double worst = 0.0;
            for (int j = 0; j < k; ++j) {
                res[j] = std::sqrt(r2[j] / std::max(mx2[j], 1e-300));
                const double rel = res[j] / std::max(std::max(std::fabs(theta[j]), floor), 1e-300);
                worst = std::max(worst, rel);
                conv[j] = rel <= kPolyLockFactor * tol;
                if (rel <= tol) ++nconv;
            }
            st.max_resid = worst;
            st.converged = nconv;
            // learned preconditioner: back to Jacobi when it stalls, i.e. the
            // geometric mean of the wanted pairs' absolute residuals
            // ||A x - theta B x|| / ||B x|| has fallen by less than 10 % over
            // the last kStallWindow iterations.  (A halving test rejects
            // every preconditioner of Jacobi's class on large meshes:
            // LOBPCG reduces the residual by 0.9 - 0.99 per iteration with
            // them at 163,842 vertices, with the network as with Jacobi, whose
            // iterations the network cut 3.6-fold there; relative residuals
            // can rise early on, while the Ritz values fall towards the
            // smallest eigenvalues.)
            if (learned_on && !no_fallback) {
                double lg = 0.0;
                for (int j = 0; j < k; ++j) lg += std::log(std::max(res[j], 1e-300));
                const double gm = std::exp(lg / k);
                const double before = hist[it % kStallWindow];  // iteration it - kStallWindow
                hist[it % kStallWindow] = gm;
                if (it >= kStallWindow && gm > 0.9 * before) {
                    learned_on = false;
                    st.fallback_used = 1;
                    if (verbose) fprintf(stderr, "[lobpcg] it %d: learned preconditioner -> Jacobi\n", it);
                }
            }
            if (verbose > 1 || (verbose && it % 50 == 0))
                fprintf(stderr, "[lobpcg] it %d conv %d/%d worst %.3e theta0 %.6e theta_k %.6e\n", it, nconv, k,
                        worst, theta[0], theta[k - 1]);
            if (nconv == k) break;
            if (phase_one) {
                // the FP32 residuals stop falling at the rounding floor
                // (~u32 ||A|| in absolute terms): hand over as soon as the
                // geometric mean of the wanted pairs' absolute residuals
                // ||A x - theta B x|| / ||B x|| is more than 0.55 times the
                // previous iteration's (~0.4 times before the floor).
                // Relative residuals would mislead early on, while the Ritz
                // values fall by orders of magnitude towards the smallest
                // eigenvalues.
                double lg = 0.0;
                for (int j = 0; j < k; ++j) lg += std::log(std::max(res[j], 1e-300));
                gm_hist.push_back(std::exp(lg / k));
                if (it >= kMinFp32Iters && gm_hist[it] > 0.55 * gm_hist[it - 1]) break;
            }
            // precondition W
            if (learned_on) {
                learned.apply(Sp() + m, ld, m);
            } else if (poly_degree > 1) {
                // soft locking of the polynomial: the wanted pairs that pass
                // the convergence test keep a plain Jacobi direction in W;
                // the others and the guard columns k..m-1, gathered in front
                // of W (ma of them, rounded up to 8), get the polynomial,
                // when that saves 8 columns or more.  The order of W's
                // columns does not matter to the Rayleigh-Ritz step.
                perm.clear();
                for (int j = 0; j < m; ++j)
                    if (j >= k || !conv[j]) perm.push_back(j);
                const int ma = std::min(m, ((int)perm.size() + 7) / 8 * 8);
                if (m - ma >= 8) {
                    for (int j = 0; j < k; ++j)
                        if (conv[j]) perm.push_back(j);
                    // staged in the context's pinned buffer and copied without a
                    // wait: the buffer is next written after the next residual
                    // read, which waits for this copy (stream order)
                    char* pin = static_cast<char*>(ctx->host_small);
                    double* nrm = reinterpret_cast<double*>(pin);
                    for (int c = 0; c < m; ++c) nrm[c] = r2[perm[c]];  // r2: ||w_c||^2
                    std::memcpy(pin + (size_t)m * 8, perm.data(), (size_t)m * 4);
                    HIPX(hipMemcpyAsync(lockbuf.p, pin, (size_t)m * 12, hipMemcpyHostToDevice, ctx->stream));
                    T* W = Sp() + m;
                    T* tmp = S2.as<T>() + m;  // S2 is free until the update below
                    gather_cols<T>(ctx, n, m, W, ld, tmp, ld,
                                   reinterpret_cast<const int32_t*>(lockbuf.as<char>() + (size_t)m * 8));
                    copy_block<T, T>(ctx, n, m, tmp, ld, W, ld);
                    cheb->apply(W, ld, ma, lockbuf.as<double>());
                    rowscale<T>(ctx, n, m - ma, W + ma, ld, dinv.as<double>());
                    st.poly_columns += ma;
                } else {
                    cheb->apply(Sp() + m, ld, m, cs1.as<double>());  // cs1: ||w_c||^2 from the test above
                    st.poly_columns += m;
                }
                st.spmm_calls += poly_degree - 1;
            } else if (precond != LH_PRECOND_NONE) {
                rowscale<T>(ctx, n, m, Sp() + m, ld, dinv.as<double>());
            }
            // [W P] orthogonal to X in the B inner product (classical Gram-Schmidt, twice)
            int nb = have_p ? 2 * m : m;
            project_out(0, m, m, nb);
            project_out(0, m, m, nb);
            // joint Cholesky-QR of [W P]; P is dropped if [W P] is numerically rank deficient
            spmm_B(Sp() + m, MSp() + m, nb);
            if (!cholqr(m, nb, true)) {
                if (nb == 2 * m) {
                    st.restarts++;
                    nb = m;
                    if (!cholqr(m, nb, true))
                        LH_THROW(LH_ERR_NUMERIC, "Cholesky-QR of the residual block failed at iteration %d", it);
                } else {
                    LH_THROW(LH_ERR_NUMERIC, "Cholesky-QR of the residual block failed at iteration %d", it);
                }
            }
            const int N = m + nb;
            {
                const double t0 = now_us();
                rr_issue(N);  // status and theta are read with the next residual norms
                rr_pending = true;
                st.t_rr_us += now_us() - t0;
            }
            // X_new = S C ; P_new = S[:, m:N] C[m:N, :]
            T* S2p = S2.as<T>();
            gemm_nn<T>(ctx, n, N, m, Sp(), ld, Cm.as<double>(), m, S2p, ld, 1.0, 0.0);
            gemm_nn<T>(ctx, n, N - m, m, Sp() + m, ld, Cm.as<double>() + (size_t)m * m, m, S2p + 2 * m, ld, 1.0,
                       0.0);
            std::swap(S, S2);
            spmm_AB(Sp(), ASp(), MSp(), m);
            have_p = true;
        }
        if (rr_pending) {  // the loop ended at max_iter: read the last Rayleigh-Ritz
            int s = 0;
            theta_new.resize(m);
            const Xfer x[2] = {{&s, status.p, sizeof(int)}, {theta_new.data(), theta_d.p, (size_t)m * 8}};
            d2h_n(ctx, x, 2);
            rr_pending = false;
            if (s) LH_THROW(LH_ERR_NUMERIC, "Rayleigh-Ritz failed at iteration %d", it - 1);
            theta.swap(theta_new);
        }
        HIPX(hipStreamSynchronize(ctx->stream));
        st.iterations = it;
        st.t_iter_us = now_us() - t_iter0;
        if (!out_vecs) {  // first phase: X stays in S
            st.t_total_us = now_us() - t_start;
            return;
        }
        // final: B-orthonormalise X and take the Ritz pairs of span(X) alone,
        // so that the returned vectors are B-orthonormal to rounding and the
        // eigenvalues are Rayleigh quotients of the returned vectors
        spmm_B(Sp(), MSp(), m);
        if (cholqr(0, m, true) && rayleigh_ritz(m)) {
            gemm_mfma_inplace<T>(ctx, n, m, Sp(), ld, Cm.as<double>(), m, 1.0, false);
        }
        for (int j = 0; j < k; ++j) out_evals[j] = theta[j];
        {
            ArenaOff keep;  // the result outlives the solver's buffers
            out_vecs->alloc(n * k * 8);
        }
        const Reorder* order = ext_ro ? ext_ro : (reordered ? ro.get() : nullptr);
        if constexpr (sizeof(T) == 8) {
            if (order) order->unpermute(ctx, Sp(), ld, out_vecs->as<double>(), k, n, k);
            else copy_block<double, double>(ctx, n, k, Sp(), ld, out_vecs->as<double>(), k);
        } else {
            LH_THROW(LH_ERR_ARG, "FP32 blocks are not returned");
        }
        HIPX(hipStreamSynchronize(ctx->stream));
        st.t_total_us = now_us() - t_start;
        if (st.converged < k) {
            st.iterations = it;
        }
    }
};

template <typename VT>
__global__ void k_densify(int64_t n, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cols,
                          const VT* __restrict__ vals, double* __restrict__ D) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int32_t s = rowptr[i]; s < rowptr[i + 1]; ++s) D[i * n + cols[s]] = (double)vals[s];
}

void densify(lh_ctx* ctx, const lh_csr* a, double* D) {
    const int64_t n = a->n();
    HIPX(hipMemsetAsync(D, 0, n * n * 8, ctx->stream));
    if (a->dtype == LH_F64)
        k_densify<double><<<nblocks(n, 256), 256, 0, ctx->stream>>>(n, a->pat->rowptr.as<int32_t>(),
                                                                   a->pat->cols.as<int32_t>(), a->vals.as<double>(), D);
    else
        k_densify<float><<<nblocks(n, 256), 256, 0, ctx->stream>>>(n, a->pat->rowptr.as<int32_t>(),
                                                                  a->pat->cols.as<int32_t>(), a->vals.as<float>(), D);
    LH_CHECK_LAUNCH();
}

void dense_eigs(lh_ctx* ctx, const lh_csr* A, const lh_csr* B, int k, double* evals, DevBuf& evecs, lh_stats& st) {
    const double t0 = now_us();
    const int n = (int)A->n();
    DevBuf GA((size_t)n * n * 8), GM((size_t)n * n * 8), C((size_t)n * k * 8), th((size_t)n * 8),
        work(rr_work_doubles(n, k) * 8), status(64);
    densify(ctx, A, GA.as<double>());
    densify(ctx, B, GM.as<double>());
    rr_solve(ctx, n, k, GA.as<double>(), GM.as<double>(), C.as<double>(), th.as<double>(), work.as<double>(),
             status.as<int>());
    int s = 0;
    d2h(ctx, &s, status.p, 4);
    if (s) LH_THROW(LH_ERR_NUMERIC, "mass matrix is not positive definite");
    d2h(ctx, evals, th.p, (size_t)k * 8);
    {
        ArenaOff keep;  // the result outlives the solver's buffers
        evecs.alloc((size_t)n * k * 8);
    }
    HIPX(hipMemcpyAsync(evecs.p, C.p, (size_t)n * k * 8, hipMemcpyDeviceToDevice, ctx->stream));
    HIPX(hipStreamSynchronize(ctx->stream));
    st.iterations = 1;
    st.converged = k;
    st.rr_calls = 1;
    st.t_total_us = now_us() - t0;
}

// the mixed-precision solve's FP32 phase hands over at this relative residual
// (or earlier, when its residuals stop falling)
constexpr double kMixedSwitchTol = 1e-5;

// block = k wanted + guard vectors, rounded to a multiple of 8, at most 512
// columns unless k itself is larger (before the 3m <= n limit of the mesh)
int block_width(int k, const lh_eig_opts* opts) {
    int m = opts && opts->block > 0 ? opts->block : ((std::max(k + 4, (k * 6 + 4) / 5) + 7) / 8) * 8;
    m = std::max(m, k);
    if (m > 512) m = std::max(k, 512);
    return m;
}

// one generalized eigenproblem on ctx's stream; throws lh_error on bad
// arguments or failures, returns the result with code LH_OK or
// LH_ERR_NOCONV (results of the last iteration)
lh_eig* eigs_impl(lh_ctx* ctx, lh_csr* A, lh_csr* B, int k, double sigma, const lh_eig_opts* opts, lh_stats& st,
                  int& code) {
    const int64_t n = A->n();
    if (B->n() != n) LH_THROW(LH_ERR_ARG, "A and B differ in size (%lld vs %lld)", (long long)n, (long long)B->n());
    if (k <= 0 || k >= n) LH_THROW(LH_ERR_ARG, "k must satisfy 0 < k < n (k=%d, n=%lld)", k, (long long)n);
    Lobpcg<double> L{};
    L.ctx = ctx;
    L.A = A;
    L.B = B;
    L.n = n;
    L.k = k;
    L.sigma = sigma;
    int m = block_width(k, opts);
    if (3 * m > n) m = std::max(k, (int)(n / 3));
    L.m = m;
    L.poly_half = (opts && opts->poly_storage == LH_STORE_F16) ||
                  ((!opts || opts->poly_storage == LH_STORE_AUTO) && m % 8 == 0 && m <= 512);
    if (L.poly_half && m % 8 != 0) LH_THROW(LH_ERR_ARG, "FP16 Chebyshev storage needs a block size divisible by 8");
    L.tol = (opts && opts->tol > 0) ? opts->tol : 1e-9;
    L.max_iter = (opts && opts->max_iter > 0) ? opts->max_iter : 20000;
    L.seed = opts ? opts->seed : 0;
    L.verbose = opts ? opts->verbose : 0;
    // preconditioner: Jacobi, accelerated by a Chebyshev polynomial on
    // [lmax / (2 d^2), lmax] whose degree d grows with sqrt(n) (the condition
    // number of a mesh Laplacian grows like n^(2/dim)) until the iteration
    // count stops falling (128 at 655k vertices: beyond it the block's own
    // gap ratio limits LOBPCG); poly_degree 1 is plain Jacobi
    int pc = opts ? opts->precond : LH_PRECOND_JACOBI;
    if (pc != LH_PRECOND_JACOBI && pc != LH_PRECOND_LEARNED && pc != LH_PRECOND_NONE && pc != LH_PRECOND_AUTO)
        LH_THROW(LH_ERR_ARG, "preconditioner %d unknown", pc);
    const int auto_deg = std::min(128, (int)std::lround(std::sqrt((double)n) / 6.3));
    if (pc == LH_PRECOND_AUTO) {
        // the learned network where the automatic polynomial would be plain
        // Jacobi (small meshes), Chebyshev-Jacobi elsewhere
        const bool small = (opts && opts->poly_degree > 0) ? opts->poly_degree == 1 : auto_deg < 4;
        const bool fp64 = !opts || opts->precision == LH_PREC_FP64;  // auto picks the network for FP64 solves only
        pc = (ctx->mpnn && small && fp64) ? LH_PRECOND_LEARNED : LH_PRECOND_JACOBI;
    }
    if (pc == LH_PRECOND_LEARNED && !ctx->mpnn)
        LH_THROW(LH_ERR_ARG, "LH_PRECOND_LEARNED needs a network on the context (lh_ctx_set_mpnn)");
    L.precond = pc;
    L.no_fallback = opts && opts->no_fallback;
    const bool jac = pc == LH_PRECOND_JACOBI;
    int deg = opts ? opts->poly_degree : 0;
    if (!jac) deg = 1;
    if (deg <= 0) deg = auto_deg < 4 ? 1 : auto_deg;
    deg = std::max(1, std::min(deg, 1024));
    L.poly_degree = deg;
    const int store = opts ? opts->poly_storage : LH_STORE_AUTO;
    if (store != LH_STORE_AUTO && store != LH_STORE_F32 && store != LH_STORE_F16)
        LH_THROW(LH_ERR_ARG, "poly_storage %d unknown", store);
    L.poly_ratio = (opts && opts->poly_ratio > 1.0) ? opts->poly_ratio : 2.0 * deg * deg;
    const int prec = opts ? opts->precision : LH_PREC_FP64;
    if (prec != LH_PREC_FP64 && prec != LH_PREC_MIXED) LH_THROW(LH_ERR_ARG, "precision mode %d unknown", prec);
    std::unique_ptr<lh_eig> e(new lh_eig());
    e->ctx = ctx;
    e->n = n;
    e->k = k;
    e->evals.assign(k, 0.0);
    if (3 * m > n || n <= 384) {
        // small problem: Rayleigh-Ritz on the whole space (S = I), i.e. the
        // dense generalized eigensolver on the GPU
        dense_eigs(ctx, A, B, k, e->evals.data(), e->evecs, L.st);
        st = L.st;
        code = LH_OK;
        return e.release();
    }
    if (prec == LH_PREC_MIXED) {
        // FP32 phase, then the FP64 solve from its X block (same order, same
        // Chebyshev preconditioner)
        Lobpcg<float> L1{};
        L1.ctx = ctx;
        L1.A = A;
        L1.B = B;
        L1.n = n;
        L1.k = k;
        L1.m = m;
        L1.sigma = sigma;
        L1.tol = std::max(L.tol, kMixedSwitchTol);
        L1.max_iter = L.max_iter;
        L1.seed = L.seed;
        L1.verbose = L.verbose;
        L1.precond = L.precond;
        L1.poly_degree = L.poly_degree;
        L1.poly_ratio = L.poly_ratio;
        L1.poly_half = L.poly_half;
        L1.phase_one = true;
        L1.run(nullptr, nullptr);
        DevBuf X0((size_t)n * m * 8);
        copy_block<float, double>(ctx, n, m, L1.Sp(), L1.ld, X0.as<double>(), m);
        L.A = L1.A;
        L.B = L1.B;
        L.ext_ro = L1.reordered ? L1.ro.get() : nullptr;
        L.cheb = L1.cheb;
        L.cheb_ready = L.poly_degree > 1;
        L.x0 = X0.as<double>();
        L.x0_ld = m;
        L.max_iter = std::max(1, L.max_iter - L1.st.iterations);
        L.run(e->evals.data(), &e->evecs);
        L.st.iterations += L1.st.iterations;
        L.st.iterations_fp32 = L1.st.iterations;
        L.st.spmm_calls += L1.st.spmm_calls;
        L.st.poly_columns += L1.st.poly_columns;
        L.st.rr_calls += L1.st.rr_calls;
        L.st.restarts += L1.st.restarts;
        L.st.t_rr_us += L1.st.t_rr_us;
        L.st.t_iter_us += L1.st.t_iter_us;
        L.st.t_setup_us += L1.st.t_setup_us;
        L.st.t_total_us += L1.st.t_total_us;
        L.st.precision_mode = LH_PREC_MIXED;
    } else {
        L.run(e->evals.data(), &e->evecs);
        L.st.precision_mode = LH_PREC_FP64;
    }
    st = L.st;
    code = (L.st.converged < k) ? LH_ERR_NOCONV : LH_OK;
    return e.release();
}

// the batched solver's workers: their own contexts (streams) and arenas
void ensure_workers(lh_ctx* ctx, int nw) {
    while ((int)ctx->workers.size() < nw) {
        lh_ctx* w = nullptr;
        const int c = lh_ctx_create(ctx->device, &w);
        if (c != LH_OK) LH_THROW(c, "could not create a worker context");
        w->tune = ctx->tune;  // the workers use the context's tuned launch shapes
        ctx->workers.push_back(w);
    }
}

// worker arena: large enough for the biggest problem's solver buffers (the
// S, A S, B S and next-S blocks dominate: 4 n (3m) doubles), plus slack for
// the preconditioner, reordering and reductions; allocations beyond it fall
// back to hipMalloc
size_t arena_bytes(int64_t nmax, int m) { return (size_t)nmax * 3 * m * 8 * 4 + (size_t)nmax * m * 8 + (256u << 20); }

struct BatchResult {
    int code = LH_OK;
    std::string err;
};

void solve_batch(lh_ctx* ctx, int nmesh, lh_csr* const* A, lh_csr* const* B, const int* k, double sigma,
                 const lh_eig_opts* opts, int nw, lh_eig** out, int* status, lh_stats* stats, BatchResult& res) {
    HIPX(hipStreamSynchronize(ctx->stream));  // inputs were made on ctx's stream
    std::vector<int> order(nmesh);
    for (int i = 0; i < nmesh; ++i) order[i] = i;
    // largest first (LPT): the long solves start early, the short ones fill in
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return A[a]->n() > A[b]->n(); });
    int64_t nmax = 0;
    int kmax = 1;
    for (int i = 0; i < nmesh; ++i) {
        nmax = std::max<int64_t>(nmax, A[i]->n());
        kmax = std::max(kmax, k[i]);
    }
    const int mmax = block_width(kmax, opts);
    nw = std::max(1, std::min(nw, nmesh));
    ensure_workers(ctx, nw);
    const size_t abytes = arena_bytes(nmax, mmax);
    for (int w = 0; w < nw; ++w) {
        lh_ctx* wc = ctx->workers[w];
        wc->mpnn = ctx->mpnn;  // the learned preconditioner's weights are shared read-only
        if (wc->arena_mem.bytes < abytes) {
            wc->arena_mem.alloc(abytes);
            wc->arena.base = wc->arena_mem.as<char>();
            wc->arena.cap = abytes;
            wc->arena.top = 0;
        }
    }
    std::atomic<int> next{0};
    std::mutex mu;
    auto work = [&](int w) {
        lh_ctx* wc = ctx->workers[w];
        (void)hipSetDevice(wc->device);
        lh_thread_arena = &wc->arena;
        for (;;) {
            const int q = next.fetch_add(1);
            if (q >= nmesh) break;
            const int i = order[q];
            const size_t mark = wc->arena.top;
            lh_stats sti{};
            int code = LH_OK;
            lh_eig* e = nullptr;
            try {
                e = eigs_impl(wc, A[i], B[i], k[i], sigma, opts, sti, code);
                if (e) e->ctx = ctx;
            } catch (const lh_error& ex) {
                code = ex.code;
                std::lock_guard<std::mutex> g(mu);
                if (res.code == LH_OK) { res.code = ex.code; res.err = "problem " + std::to_string(i) + ": " + ex.msg; }
            } catch (const std::exception& ex) {
                code = LH_ERR_HIP;
                std::lock_guard<std::mutex> g(mu);
                if (res.code == LH_OK) { res.code = LH_ERR_HIP; res.err = ex.what(); }
            }
            (void)hipStreamSynchronize(wc->stream);
            wc->arena.top = mark;  // every buffer of the problem is released
            out[i] = e;
            if (status) status[i] = code;
            if (stats) stats[i] = sti;
            if (code == LH_ERR_NOCONV) {
                std::lock_guard<std::mutex> g(mu);
                if (res.code == LH_OK) { res.code = LH_ERR_NOCONV; res.err = "problem " + std::to_string(i) + " did not converge"; }
            }
        }
        lh_thread_arena = nullptr;
    };
    std::vector<std::thread> th;
    for (int w = 1; w < nw; ++w) th.emplace_back(work, w);
    work(0);
    for (auto& t : th) t.join();
}

}  // namespace

extern "C" {

int lh_eigs(lh_ctx* ctx, lh_csr* A, lh_csr* B, int k, double sigma, const lh_eig_opts* opts, lh_eig** out,
            lh_stats* stats) {
    if (!ctx || !A || !B || !out) return LH_ERR_ARG;
    *out = nullptr;
    LH_API_BEGIN(ctx)
    HIPX(hipSetDevice(ctx->device));
    lh_stats st{};
    int code = LH_OK;
    *out = eigs_impl(ctx, A, B, k, sigma, opts, st, code);
    if (stats) *stats = st;
    if (code == LH_ERR_NOCONV) {
        ctx->last_error = "LOBPCG did not converge within max_iter";
        return LH_ERR_NOCONV;
    }
    LH_API_END(ctx)
}

int lh_eigs_batched(lh_ctx* ctx, int nmesh, lh_csr* const* A, lh_csr* const* B, const int* k, double sigma,
                    const lh_eig_opts* opts, int workers, lh_eig** out, int* status, lh_stats* stats) {
    if (!ctx || nmesh < 0 || (nmesh > 0 && (!A || !B || !k || !out))) return LH_ERR_ARG;
    for (int i = 0; i < nmesh; ++i) out[i] = nullptr;
    if (nmesh == 0) return LH_OK;
    LH_API_BEGIN(ctx)
    HIPX(hipSetDevice(ctx->device));
    for (int i = 0; i < nmesh; ++i)
        if (!A[i] || !B[i]) LH_THROW(LH_ERR_ARG, "problem %d has no matrix", i);
    if (workers < 0 || workers > 64) LH_THROW(LH_ERR_ARG, "workers must be in [0, 64], got %d", workers);
    const int nw = workers > 0 ? workers : (ctx->tuned_workers > 0 ? ctx->tuned_workers : 8);
    BatchResult res;
    solve_batch(ctx, nmesh, A, B, k, sigma, opts, nw, out, status, stats, res);
    if (res.code != LH_OK) {
        ctx->last_error = res.err;
        return res.code;
    }
    LH_API_END(ctx)
}

namespace {

// the ELL SpMM pair's workgroup size for one problem shape: 64, 128 and 256
// threads timed (HIP events, 20 launches each) on the solver's block width,
// the fastest kept in the context's table (a shape already there is kept)
void tune_spmm_shape(lh_ctx* ctx, const lh_csr* A, const lh_csr* B, int k, const lh_eig_opts* opts) {
    const int64_t n = A->n();
    const int m = block_width(k, opts);
    if (3 * m > n || n <= 384 || A->dtype != LH_F64) return;  // the dense path: no SpMM
    if (tuned_spmm_block(ctx, n, A->nnz())) return;
    Ell e;
    e.build(ctx, A, B);
    if (!e.W) return;
    DevBuf X((size_t)n * m * 8), YA((size_t)n * m * 8), YB((size_t)n * m * 8);
    memset_d(ctx, X.p, 0, (size_t)n * m * 8);
    double* yb = e.two ? YB.as<double>() : nullptr;
    static const int cand[3] = {64, 128, 256};
    hipEvent_t e0, e1;
    HIPX(hipEventCreate(&e0));
    HIPX(hipEventCreate(&e1));
    double best = 1e300;
    int bb = 0;
    for (int c = 0; c < 3; ++c) {
        if (!spmm_ell<double, double>(ctx, e, X.as<double>(), YA.as<double>(), yb, m, m, m, 0, cand[c])) break;
        HIPX(hipEventRecord(e0, ctx->stream));
        for (int it = 0; it < 20; ++it)
            spmm_ell<double, double>(ctx, e, X.as<double>(), YA.as<double>(), yb, m, m, m, 0, cand[c]);
        HIPX(hipEventRecord(e1, ctx->stream));
        HIPX(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPX(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
            best = ms;
            bb = cand[c];
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!bb) return;
    std::lock_guard<std::mutex> lk(ctx->tune->mu);
    ctx->tune->spmm.push_back(SpmmTune{n, A->nnz(), k, m, bb});
}

}  // namespace

int lh_eigs_autotune(lh_ctx* ctx, int nmesh, lh_csr* const* A, lh_csr* const* B, const int* k, double sigma,
                     const lh_eig_opts* opts, int* best_workers, double* seconds) {
    if (!ctx || nmesh <= 0 || !A || !B || !k) return LH_ERR_ARG;
    LH_API_BEGIN(ctx)
    HIPX(hipSetDevice(ctx->device));
    static const int cand[5] = {1, 2, 4, 8, 16};
    double best = 1e300;
    int bw = 1;
    std::vector<lh_eig*> out(nmesh, nullptr);
    for (int c = 0; c < 5; ++c) {
        BatchResult res;
        const double t0 = now_us();
        solve_batch(ctx, nmesh, A, B, k, sigma, opts, cand[c], out.data(), nullptr, nullptr, res);
        const double dt = (now_us() - t0) * 1e-6;
        for (auto& e : out) { lh_eig_free(e); e = nullptr; }
        if (res.code != LH_OK && res.code != LH_ERR_NOCONV) {
            ctx->last_error = res.err;
            return res.code;
        }
        if (seconds) seconds[c] = dt;
        if (dt < best) { best = dt; bw = cand[c]; }
    }
    ctx->tuned_workers = bw;
    if (best_workers) *best_workers = bw;
    // then the SpMM's workgroup size per problem shape
    for (int i = 0; i < nmesh; ++i) tune_spmm_shape(ctx, A[i], B[i], k[i], opts);
    LH_API_END(ctx)
}

int lh_tune_count(lh_ctx* ctx, int* count) {
    if (!ctx || !count) return LH_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->tune->mu);
    *count = (int)ctx->tune->spmm.size();
    return LH_OK;
}

int lh_tune_get(lh_ctx* ctx, int i, lh_tune_entry* out) {
    if (!ctx || !out) return LH_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->tune->mu);
    if (i < 0 || i >= (int)ctx->tune->spmm.size()) return LH_ERR_ARG;
    const SpmmTune& t = ctx->tune->spmm[i];
    *out = lh_tune_entry{t.n, t.nnz, t.k, t.ncol, t.block, 0};
    return LH_OK;
}

int lh_tune_set(lh_ctx* ctx, const lh_tune_entry* e) {
    if (!ctx || !e || e->n <= 0 || e->nnz < 0 || (e->block != 64 && e->block != 128 && e->block != 256))
        return LH_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->tune->mu);
    for (SpmmTune& t : ctx->tune->spmm)
        if (t.n == e->n && t.nnz == e->nnz) {
            t = SpmmTune{e->n, e->nnz, e->k, e->ncol, e->block};
            return LH_OK;
        }
    ctx->tune->spmm.push_back(SpmmTune{e->n, e->nnz, e->k, e->ncol, e->block});
    return LH_OK;
}

int lh_ctx_get_workers(lh_ctx* ctx, int* workers) {
    if (!ctx || !workers) return LH_ERR_ARG;
    *workers = ctx->tuned_workers;
    return LH_OK;
}

int lh_ctx_set_workers(lh_ctx* ctx, int workers) {
    if (!ctx || workers < 0 || workers > 64) return LH_ERR_ARG;
    ctx->tuned_workers = workers;
    return LH_OK;
}

int lh_device_arch(lh_ctx* ctx, char* buf, int len) {
    if (!ctx || !buf || len <= 0) return LH_ERR_ARG;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->device) != hipSuccess) return LH_ERR_HIP;
    std::snprintf(buf, (size_t)len, "%s", prop.gcnArchName);
    return LH_OK;
}

int lh_eig_download(lh_eig* e, double* evals, double* evecs) {
    if (!e) return LH_ERR_ARG;
    lh_ctx* ctx = e->ctx;
    LH_API_BEGIN(ctx)
    HIPX(hipSetDevice(ctx->device));
    if (evals) for (int j = 0; j < e->k; ++j) evals[j] = e->evals[j];
    if (evecs) {
        // device row-major n x k == host column-major k x n transposed; the
        // caller wants evecs[:, j] contiguous (Fortran order), so transpose.
        std::vector<double> h((size_t)e->n * e->k);
        d2h(ctx, h.data(), e->evecs.p, h.size() * 8);
        for (int64_t i = 0; i < e->n; ++i)
            for (int j = 0; j < e->k; ++j) evecs[(int64_t)j * e->n + i] = h[i * e->k + j];
    }
    LH_API_END(ctx)
}

void lh_eig_free(lh_eig* e) {
    if (!e) return;
    (void)hipSetDevice(e->ctx->device);
    (void)hipStreamSynchronize(e->ctx->stream);
    delete e;
}

}  // extern "C"
