// lsvd_bss.cpp -- LieSvdBss::separate (src/lie_svd_bss.rs:58-101) for a batch
// of same-shape observation blocks: the covariances of all blocks go through
// LieSvdSmall::solve in one batched call of the small-SVD kernels and their
// lagged families through one batched Phase-JADE call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "lsvd_api_internal.hpp"
#include "lsvd_bss.hpp"

using namespace lsvd;

namespace {

// temporaries come from the handle's workspace (lsvd_workspace.hpp)
using DMem = lsvd::WsMem;

// Runs the joint diagonalizer on device buffers whatever the caller's pointer
// mode, restoring it afterwards.
struct DeviceMode {
    lsvd_handle_t h;
    int saved;
    explicit DeviceMode(lsvd_handle_t hh) : h(hh), saved(hh->ptr_mode) { h->ptr_mode = LSVD_PTR_DEVICE; }
    ~DeviceMode() { h->ptr_mode = saved; }
};

}  // namespace

#define BCHK(expr)                                          \
    do {                                                    \
        hipError_t e_ = (expr);                             \
        if (e_ != hipSuccess) return hip_fail(h, e_, #expr); \
    } while (0)

extern "C" {

int lsvd_bss_separate_f64(lsvd_handle_t h, int count, int channels, int samples, const double* x, int nlags,
                          const int* lags, double eps, const lsvd_joint_params_t* params, double* unmixing,
                          double* separated, double* coherence, double* whitening, double* whitened,
                          lsvd_bss_trace_t* trace) {
    if (!h) return LSVD_E_ARG;
    if (count < 0 || nlags < 0) return set_error(h, LSVD_E_SHAPE, "negative size");
    if (channels <= 0 || samples <= 1) return set_error(h, LSVD_E_SHAPE, "PhaseBSS: empty observations");
    if (channels > kBssMaxChannels)
        return set_error(h, LSVD_E_LIMIT, "PhaseBSS: channels > " + std::to_string(kBssMaxChannels));
    if (count == 0) return LSVD_OK;
    if (!x || !unmixing || !separated || !coherence || (nlags > 0 && !lags))
        return set_error(h, LSVD_E_ARG, "null pointer");
    BCHK(hipSetDevice(h->device));
    hipStream_t st = h->stream;
    const bool host = h->ptr_mode == LSVD_PTR_HOST;
    const int ch = channels, S = samples;
    const size_t blk = (size_t)ch * S, cc = (size_t)ch * ch;
    // the usable lags (:181-183); none: the covariance itself (:192-194)
    std::vector<int> lag;
    std::vector<double> denom;
    for (int l = 0; l < nlags; ++l)
        if (lags[l] > 0 && lags[l] < S) {
            lag.push_back(lags[l]);
            denom.push_back((double)(S - lags[l]));
        }
    if (lag.empty()) {
        lag.push_back(0);
        denom.push_back((double)S);
    }
    const int K = (int)lag.size();
    const int ns_cov = bss_splits(count, ch, S, 1), ns_lag = bss_splits(count, ch, S, K);
    lsvd::WsScope ws_scope(h);
    DMem bx, bmean, bpart, bcov, bu, bs, bvt, bw, bz, bfloor, bfam, bv, bdiag, bun, bsep, bcoh, blag, bden;
    const double* dx = x;
    if (host) {
        BCHK(bx.alloc(sizeof(double) * blk * count));
        BCHK(hipMemcpyAsync(bx.p, x, sizeof(double) * blk * count, hipMemcpyHostToDevice, st));
        dx = (const double*)bx.p;
    }
    const int zero = 0;
    const double fS = (double)S;
    BCHK(blag.alloc(sizeof(int) * (K + 1)));
    BCHK(bden.alloc(sizeof(double) * (K + 1)));
    BCHK(hipMemcpyAsync(blag.p, lag.data(), sizeof(int) * K, hipMemcpyHostToDevice, st));
    BCHK(hipMemcpyAsync((int*)blag.p + K, &zero, sizeof(int), hipMemcpyHostToDevice, st));
    BCHK(hipMemcpyAsync(bden.p, denom.data(), sizeof(double) * K, hipMemcpyHostToDevice, st));
    BCHK(hipMemcpyAsync((double*)bden.p + K, &fS, sizeof(double), hipMemcpyHostToDevice, st));
    const int* d_lag = (const int*)blag.p;
    const double* d_den = (const double*)bden.p;
    // centered covariance (:60-63)
    BCHK(bmean.alloc(sizeof(double) * ch * count));
    BCHK(bpart.alloc(sizeof(double) * cc * count * std::max(ns_cov, K * ns_lag)));
    BCHK(bcov.alloc(sizeof(double) * cc * count));
    double* dmean = (double*)bmean.p;
    BCHK(bss_means(dx, count, ch, S, dmean, st));
    BCHK(bss_cross(dx, dmean, count, ch, S, d_lag + K, 1, ns_cov, (double*)bpart.p, st));
    BCHK(bss_reduce((const double*)bpart.p, count, ch, 1, ns_cov, d_den + K, false, (double*)bcov.p, st));
    // LieSvdSmall::solve of every covariance (:65)
    BCHK(bu.alloc(sizeof(double) * cc * count));
    BCHK(bvt.alloc(sizeof(double) * cc * count));
    BCHK(bs.alloc(sizeof(double) * ch * count));
    std::vector<int32_t> nv(count, ch);
    std::vector<int64_t> moff(count), soff(count);
    for (int p = 0; p < count; ++p) {
        moff[p] = (int64_t)p * (int64_t)cc;
        soff[p] = (int64_t)p * ch;
    }
    int rc = svd_device_small(h, count, nv.data(), moff.data(), soff.data(), LSVD_PREC_F64, (const double*)bcov.p,
                              (double*)bu.p, (double*)bs.p, (double*)bvt.p, nullptr);
    if (rc != LSVD_OK) return rc;
    // whitening (:66-74)
    double* dW = whitening;
    double* dZ = whitened;
    if (host || !whitening) {
        BCHK(bw.alloc(sizeof(double) * cc * count));
        dW = (double*)bw.p;
    }
    if (host || !whitened) {
        BCHK(bz.alloc(sizeof(double) * blk * count));
        dZ = (double*)bz.p;
    }
    BCHK(bfloor.alloc(sizeof(double) * count));
    BCHK(bss_whitening((const double*)bu.p, (const double*)bs.p, count, ch, eps, dW, (double*)bfloor.p, st));
    BCHK(bss_apply(dW, false, dx, dmean, count, ch, S, dZ, st));
    // lagged family (:75, :173-197) and Phase-JADE (:76-80)
    BCHK(bfam.alloc(sizeof(double) * cc * K * count));
    BCHK(bss_cross(dZ, nullptr, count, ch, S, d_lag, K, ns_lag, (double*)bpart.p, st));
    BCHK(bss_reduce((const double*)bpart.p, count, ch, K, ns_lag, d_den, true, (double*)bfam.p, st));
    BCHK(bv.alloc(sizeof(double) * cc * count));
    BCHK(bdiag.alloc(sizeof(double) * ch * K * count));
    std::vector<lsvd_joint_trace_t> jt(count);
    {
        DeviceMode dm(h);
        rc = lsvd_joint_diag_sym_f64(h, count, K, ch, (const double*)bfam.p, params, (double*)bv.p,
                                     (double*)bdiag.p, jt.data());
    }
    if (rc != LSVD_OK) return rc;
    // separated = V^T whitened, unmixing = V^T W, coherence (:81-84)
    double* dsep = separated;
    double* dun = unmixing;
    double* dcoh = coherence;
    if (host) {
        BCHK(bsep.alloc(sizeof(double) * blk * count));
        BCHK(bun.alloc(sizeof(double) * cc * count));
        BCHK(bcoh.alloc(sizeof(double) * ch * count));
        dsep = (double*)bsep.p;
        dun = (double*)bun.p;
        dcoh = (double*)bcoh.p;
    }
    BCHK(bss_apply((const double*)bv.p, true, dZ, nullptr, count, ch, S, dsep, st));
    BCHK(bss_unmixing((const double*)bv.p, dW, count, ch, dun, st));
    BCHK(bss_coherence(dsep, count, ch, S, dcoh, st));
    if (host) {
        BCHK(hipMemcpyAsync(separated, dsep, sizeof(double) * blk * count, hipMemcpyDeviceToHost, st));
        BCHK(hipMemcpyAsync(unmixing, dun, sizeof(double) * cc * count, hipMemcpyDeviceToHost, st));
        BCHK(hipMemcpyAsync(coherence, dcoh, sizeof(double) * ch * count, hipMemcpyDeviceToHost, st));
        if (whitening) BCHK(hipMemcpyAsync(whitening, dW, sizeof(double) * cc * count, hipMemcpyDeviceToHost, st));
        if (whitened) BCHK(hipMemcpyAsync(whitened, dZ, sizeof(double) * blk * count, hipMemcpyDeviceToHost, st));
    }
    if (trace) {
        std::vector<double> fl(count), co((size_t)ch * count);
        BCHK(hipMemcpyAsync(fl.data(), bfloor.p, sizeof(double) * count, hipMemcpyDeviceToHost, st));
        BCHK(hipMemcpyAsync(co.data(), dcoh, sizeof(double) * ch * count, hipMemcpyDeviceToHost, st));
        BCHK(hipStreamSynchronize(st));
        for (int p = 0; p < count; ++p) {
            double m = 0.0;
            for (int i = 0; i < ch; ++i) m += co[(size_t)p * ch + i];
            trace[p].channels = ch;
            trace[p].samples = S;
            trace[p].covariance_floor = fl[p];
            trace[p].joint_initial_offdiag = jt[p].initial_offdiag;
            trace[p].joint_final_offdiag = jt[p].final_offdiag;
            trace[p].joint_rotations = jt[p].rotations;
            trace[p].mean_channel_coherence = m / ch;
        }
    }
    BCHK(hipStreamSynchronize(st));
    return LSVD_OK;
}

}  // extern "C"

#undef BCHK
