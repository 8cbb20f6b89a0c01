// lsvd_polar.cpp -- host driver of the Newton-Schulz polar decomposition
// for large matrices (newton_schulz_polar, src/lie_svd_small.rs:158-183).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "lsvd_api_internal.hpp"
#include "lsvd_polar.hpp"

using namespace lsvd;

namespace {

// temporaries come from the handle's workspace (lsvd_workspace.hpp)
using Buf = lsvd::WsMem;

// Pinned host slots and one event per slot for the stop-test reads of the
// plain FP64 steps.
struct StepSlots {
    double* host = nullptr;  // 2 slots x 4 doubles: ||dX||^2, unused, bad flag (as int), unused
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipError_t create() {
        hipError_t e = hipHostMalloc((void**)&host, sizeof(double) * 8, hipHostMallocDefault);
        for (int i = 0; i < 2 && e == hipSuccess; ++i) e = hipEventCreateWithFlags(&ev[i], hipEventDisableTiming);
        return e;
    }
    ~StepSlots() {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        if (host) (void)hipHostFree(host);
    }
};

}  // namespace

namespace lsvd {

// The iteration on device buffers: dA (m x n) in, dQ (m x n) out.
int polar_run(lsvd_handle_t h, const double* dA, int m, int n, int precision, double* dQ, lsvd_polar_info_t* info) {
    hipStream_t st = h->stream;
    const size_t mn = (size_t)m * n, nn = (size_t)n * n;
    const bool lowp = precision == LSVD_PREC_F32 || precision == LSVD_PREC_BF16;
    const int lp = precision == LSVD_PREC_F32 ? kPolF32 : kPolBF16;
    lsvd_polar_info_t inf{};
    lsvd::WsScope ws_scope(h);
    Buf bX, bX2, bX3, bM, bXf, bXf2, bMf, bS;
#define PCHK(expr)                                          \
    do {                                                    \
        hipError_t e_ = (expr);                             \
        if (e_ != hipSuccess) return hip_fail(h, e_, #expr); \
    } while (0)

    PCHK(bX.alloc(sizeof(double) * mn));
    PCHK(bX2.alloc(sizeof(double) * mn));
    PCHK(bM.alloc(sizeof(double) * nn));
    PCHK(bS.alloc(sizeof(double) * (4 + kSumsqBlocks + 8)));
    if (lowp) {
        PCHK(bXf.alloc(sizeof(float) * mn));
        PCHK(bXf2.alloc(sizeof(float) * mn));
        PCHK(bMf.alloc(sizeof(float) * nn));
    }
    double* S = (double*)bS.p;  // [0] ||A||^2, [1] ||dX||^2, [2] orth^2, [3] bad flag (as int)
    int* bad = (int*)(S + 3);
    double hs[4];
    PCHK(polar_zero(S, 4, st));
    PCHK(polar_sumsq(dA, mn, S, S + 4, st));
    const double rootn = std::sqrt((double)n);
    double delta = 0.0;
    int it_low = 0, it64 = 0;
    bool diverged = false;
    if (lowp) {
        PCHK(polar_scale(dA, mn, S, bXf.p, true, st));
        void* X = bXf.p;
        void* Xn = bXf2.p;
        // The low-precision phase runs until ||X^T X - I||_F (monotone in
        // exact arithmetic) reaches its rounding floor or stops improving.
        // Every direction still escaping from a small singular value adds
        // about 1 to it, so the stall test only applies below 0.5, where
        // all directions are in the quadratic regime.
        const double floor_tol = (lp == kPolBF16 ? 1e-3 : 1e-6) * rootn;
        double best = 1e300;
        int stall = 0;
        NsSchedule sched;  // scaled steps first; the stop tests apply to plain steps
        sched.init(true);
        for (int it = 0; it < 120; ++it) {
            PCHK(polar_zero(S + 1, 3, st));
            PCHK(ns_gram(lp, m, n, X, bMf.p, S + 2, st, sched.alpha()));
            PCHK(ns_update(lp, m, n, X, bMf.p, Xn, S + 1, bad, st, sched.c()));
            PCHK(hipMemcpyAsync(hs, S, sizeof(hs), hipMemcpyDeviceToHost, st));
            PCHK(hipStreamSynchronize(st));
            ++it_low;
            delta = std::sqrt(hs[1]);
            int b = 0;
            std::memcpy(&b, &hs[3], sizeof(int));
            if (b || !(delta <= 1e6 * rootn)) {  // non-finite or diverging: redo in FP64
                diverged = true;
                break;
            }
            std::swap(X, Xn);
            const double orth = std::sqrt(std::max(hs[2], 0.0));  // of the iterate before this step
            const bool plain = sched.plain();
            sched.after(hs[2], delta);
            if (!plain) continue;
            if (orth < floor_tol) break;
            if (orth < 0.99 * best) {
                best = orth;
                stall = 0;
            } else if (orth < 0.5 && ++stall >= 3) {
                break;
            }
        }
        if (diverged) {
            PCHK(polar_scale(dA, mn, S, bX.p, false, st));
        } else {
            PCHK(polar_widen((const float*)X, mn, (double*)bX.p, st));
        }
    } else {
        PCHK(polar_scale(dA, mn, S, bX.p, false, st));
    }
    double* X = (double*)bX.p;
    double* Xn = (double*)bX2.p;
    bool converged = false, bad64 = false;
    const int budget = 120;
    NsSchedule sched;  // plain steps unless LSVD_PREC_F64_SCALED
    sched.init(precision == LSVD_PREC_F64_SCALED);
    // Plain steps on a matrix up to about 1600^2, not after a low-precision
    // phase: one step ahead of the stop test.  Step k + 1 is queued before
    // the host reads step k's ||dX||, so the GPU does not idle through that
    // read (~35 us a step); the step queued after the last one is wasted,
    // which costs less than the reads save only while a step is short and
    // the iteration long.  Iterate X_k lives in buffer k % 3, so the queued
    // step (X_{k+1} -> X_{k+2}) leaves X_{k+1}, the result when step k meets
    // the stop test, in place; step k uses slot k % 2 of the device and host
    // stop-test slots.  The iterates and the step count are those of the
    // one-step-at-a-time loop.
    const bool ahead = !sched.enabled && !lowp && (double)m * n * n <= 4e9;
    if (ahead) {
        PCHK(bX3.alloc(sizeof(double) * mn));
        StepSlots ss;
        PCHK(ss.create());
        double* bufs[3] = {X, Xn, (double*)bX3.p};
        double* T = S + 4 + kSumsqBlocks;  // device slots, 4 doubles each
        auto launch = [&](int k) -> hipError_t {
            double* t = T + 4 * (k & 1);
            hipError_t e = polar_zero(t, 4, st);
            if (e == hipSuccess) e = ns_gram(kPolF64, m, n, bufs[k % 3], bM.p, nullptr, st, sched.alpha());
            if (e == hipSuccess)
                e = ns_update(kPolF64, m, n, bufs[k % 3], bM.p, bufs[(k + 1) % 3], t, (int*)(t + 2), st, sched.c());
            if (e == hipSuccess)
                e = hipMemcpyAsync(ss.host + 4 * (k & 1), t, sizeof(double) * 4, hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipEventRecord(ss.ev[k & 1], st);
            return e;
        };
        PCHK(launch(0));
        int k = 0;
        X = bufs[budget % 3];  // the iterate when the budget runs out
        for (; k < budget; ++k) {
            if (k + 1 < budget) PCHK(launch(k + 1));
            PCHK(hipEventSynchronize(ss.ev[k & 1]));
            const double* hsk = ss.host + 4 * (k & 1);
            ++it64;
            delta = std::sqrt(hsk[0]);
            int b = 0;
            std::memcpy(&b, &hsk[2], sizeof(int));
            if (b || delta > 1e6 * rootn) {  // (:169-171)
                bad64 = true;
                X = bufs[k % 3];
                break;
            }
            if (delta < 1e-13 * rootn) {
                converged = true;
                X = bufs[(k + 1) % 3];
                break;
            }
        }
        PCHK(hipStreamSynchronize(st));  // the queued step's copy into ss.host has landed
    } else {
        for (int it = 0; it < budget; ++it) {
            PCHK(polar_zero(S + 1, 3, st));
            PCHK(ns_gram(kPolF64, m, n, X, bM.p, sched.enabled ? S + 2 : nullptr, st, sched.alpha()));
            PCHK(ns_update(kPolF64, m, n, X, bM.p, Xn, S + 1, bad, st, sched.c()));
            PCHK(hipMemcpyAsync(hs, S, sizeof(hs), hipMemcpyDeviceToHost, st));
            PCHK(hipStreamSynchronize(st));
            ++it64;
            delta = std::sqrt(hs[1]);
            int b = 0;
            std::memcpy(&b, &hs[3], sizeof(int));
            if (b || delta > 1e6 * rootn) {  // (:169-171); the Gram-Schmidt fallback is not run at this size
                bad64 = true;
                break;
            }
            std::swap(X, Xn);
            const bool plain = sched.plain();
            sched.after(hs[2], delta);
            if (plain && delta < 1e-13 * rootn) {
                converged = true;
                break;
            }
        }
    }
    PCHK(polar_zero(S + 2, 1, st));
    PCHK(ns_orth(m, n, X, S + 2, st));
    // A-dependent check: Q is A's polar factor iff Q^T A is symmetric
    // (positive semidefinite), so ||Q^T A - A^T Q||_F / ||A||_F measures
    // how far Q is from it in the directions orthogonality cannot see
    // (one FP64 n x n x m product into M)
    PCHK(gemm_f64(true, n, n, m, X, dA, (double*)bM.p, 1.0, st));
    PCHK(polar_skew_sumsq((const double*)bM.p, n, S + 1, S + 4, st));
    PCHK(hipMemcpyAsync(hs, S, sizeof(hs), hipMemcpyDeviceToHost, st));
    PCHK(hipStreamSynchronize(st));
    inf.orth_err = std::sqrt(std::max(hs[2], 0.0));
    inf.skew_err = std::sqrt(std::max(hs[1], 0.0)) / std::max(std::sqrt(hs[0]), 1e-300);
    // the reference's trust test (:174, :178), plus the skew test at the
    // same threshold: a factor that is orthogonal but rotated away from A's
    // polar factor in its small-sigma directions is not trusted
    inf.polar_ok = (!bad64 && inf.orth_err <= 1e-8 * rootn && inf.skew_err <= 1e-8 * rootn) ? 1 : 0;
    (void)converged;
    inf.lowp_iters = it_low;
    inf.f64_iters = it64;
    inf.last_delta = delta;
    PCHK(hipMemcpyAsync(dQ, X, sizeof(double) * mn, hipMemcpyDeviceToDevice, st));
    PCHK(hipStreamSynchronize(st));
#undef PCHK
    if (info) *info = inf;
    return LSVD_OK;
}

}  // namespace lsvd

extern "C" {

// NS polar factor of a tall or square m x n matrix (m >= n).
//   LSVD_PREC_F64 : the reference iteration in FP64 (X0 = A/||A||_F, X <- X(3I - X^T X)/2,
//                   stop at ||dX||_F < 1e-13 sqrt(n), polar_ok = ||Q^T Q - I||_F <= 1e-8 sqrt(n)).
//   LSVD_PREC_F64_SCALED : FP64 with the scaled step schedule (NsSchedule), same limit.
//   LSVD_PREC_F32 / LSVD_PREC_BF16 : the same iteration with FP32 / BF16 MFMA products
//                   (FP32 accumulation) until it stalls at that precision, then FP64 refinement
//                   iterations with the FP64 stopping rule.
int lsvd_polar_f64(lsvd_handle_t h, const double* a, int m, int n, int precision, double* q,
                   lsvd_polar_info_t* info) {
    if (!h) return LSVD_E_ARG;
    if (m < n || n < 0) return set_error(h, LSVD_E_SHAPE, "lsvd_polar_f64: needs m >= n");
    if (precision != LSVD_PREC_F64 && precision != LSVD_PREC_F32 && precision != LSVD_PREC_BF16 &&
        precision != LSVD_PREC_F64_SCALED)
        return set_error(h, LSVD_E_ARG, "unknown precision");
    lsvd_polar_info_t inf{};
    if (n == 0) {
        inf.polar_ok = 1;
        if (info) *info = inf;
        return LSVD_OK;
    }
    if (!a || !q) return set_error(h, LSVD_E_ARG, "null pointer");
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(h, e, "hipSetDevice");
    hipStream_t st = h->stream;
    const size_t mn = (size_t)m * n;
#define PCHK(expr)                                          \
    do {                                                    \
        hipError_t e_ = (expr);                             \
        if (e_ != hipSuccess) return hip_fail(h, e_, #expr); \
    } while (0)
    if (h->ptr_mode != LSVD_PTR_HOST) return polar_run(h, a, m, n, precision, q, info);
    lsvd::WsScope ws_scope(h);
    Buf bA, bQ;
    PCHK(bA.alloc(sizeof(double) * mn));
    PCHK(bQ.alloc(sizeof(double) * mn));
    PCHK(hipMemcpyAsync(bA.p, a, sizeof(double) * mn, hipMemcpyHostToDevice, st));
    int rc = polar_run(h, (const double*)bA.p, m, n, precision, (double*)bQ.p, info);
    if (rc != LSVD_OK) return rc;
    PCHK(hipMemcpyAsync(q, bQ.p, sizeof(double) * mn, hipMemcpyDeviceToHost, st));
    PCHK(hipStreamSynchronize(st));
#undef PCHK
    return LSVD_OK;
}


// Plain FP64 product C = op(A) B (diagnostic; used by the GEMM parity tests).
int lsvd_gemm_f64(lsvd_handle_t h, int trans_a, int m, int n, int k, const double* a, const double* b, double* c) {
    if (!h || !a || !b || !c) return LSVD_E_ARG;
    if (m <= 0 || n <= 0 || k <= 0) return set_error(h, LSVD_E_SHAPE, "empty product");
    if (h->ptr_mode != LSVD_PTR_DEVICE) return set_error(h, LSVD_E_ARG, "lsvd_gemm_f64 needs device pointers");
    hipError_t e = gemm_f64(trans_a != 0, m, n, k, a, b, c, 1.0, h->stream);
    return e == hipSuccess ? LSVD_OK : hip_fail(h, e, "gemm_f64");
}

// Low-precision product (prec LSVD_PREC_F32 or LSVD_PREC_BF16), FP32 in/out.
int lsvd_gemm_lowp(lsvd_handle_t h, int precision, int trans_a, int m, int n, int k, const float* a, const float* b,
                   float* c) {
    if (!h || !a || !b || !c) return LSVD_E_ARG;
    if (m <= 0 || n <= 0 || k <= 0) return set_error(h, LSVD_E_SHAPE, "empty product");
    if (precision != LSVD_PREC_F32 && precision != LSVD_PREC_BF16) return set_error(h, LSVD_E_ARG, "precision");
    if (h->ptr_mode != LSVD_PTR_DEVICE) return set_error(h, LSVD_E_ARG, "lsvd_gemm_lowp needs device pointers");
    hipError_t e = gemm_lowp(precision == LSVD_PREC_F32 ? kPolF32 : kPolBF16, trans_a != 0, m, n, k, a, b, c,
                             h->stream);
    return e == hipSuccess ? LSVD_OK : hip_fail(h, e, "gemm_lowp");
}

}  // extern "C"
