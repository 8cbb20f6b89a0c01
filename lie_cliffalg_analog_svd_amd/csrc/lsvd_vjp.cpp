// lsvd_vjp.cpp -- host side of the SVD / polar adjoint entry points.
#include <hip/hip_runtime.h>

#include <vector>

#include "lsvd_api_internal.hpp"
#include "lsvd_vjp.hpp"

using namespace lsvd;

namespace {
// temporaries come from the handle's workspace (lsvd_workspace.hpp)
using Buf = lsvd::WsMem;
}  // namespace

static int vjp_common(lsvd_handle_t h, int count, int m, int n, const double* a, const double* u, const double* s,
                      const double* vt, const double* gu, const double* gs, const double* gvt, const double* qbar,
                      double guard, double* grad) {
    if (!h) return LSVD_E_ARG;
    if (count < 0 || m < n || n < 1) return set_error(h, LSVD_E_SHAPE, "vjp: needs count >= 0 and m >= n >= 1");
    if (count == 0) return LSVD_OK;
    if (!grad) return set_error(h, LSVD_E_ARG, "null grad");
    const bool have_fwd = u && s && vt;
    if (!have_fwd && !a) return set_error(h, LSVD_E_ARG, "need A or the forward outputs");
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(h, e, "hipSetDevice");
    hipStream_t st = h->stream;
    const bool host = h->ptr_mode == LSVD_PTR_HOST;
    const size_t mn = (size_t)m * n, nn = (size_t)n * n, cnt = (size_t)count;
#define VCHK(expr)                                          \
    do {                                                    \
        hipError_t e_ = (expr);                             \
        if (e_ != hipSuccess) return hip_fail(h, e_, #expr); \
    } while (0)
    lsvd::WsScope ws_scope(h);
    Buf bA, bU, bS, bVt, bgU, bgS, bgVt, bQ, bG, bW;
    auto stage = [&](Buf& b, const double* src, size_t len, const double** dst) -> hipError_t {
        if (!src) {
            *dst = nullptr;
            return hipSuccess;
        }
        if (!host) {
            *dst = src;
            return hipSuccess;
        }
        hipError_t ee = b.alloc(sizeof(double) * len);
        if (ee == hipSuccess) ee = hipMemcpyAsync(b.p, src, sizeof(double) * len, hipMemcpyHostToDevice, st);
        *dst = (const double*)b.p;
        return ee;
    };
    const double *dA, *dU, *dS, *dVt, *dgU, *dgS, *dgVt, *dQ;
    VCHK(stage(bA, a, cnt * mn, &dA));
    VCHK(stage(bgU, gu, cnt * mn, &dgU));
    VCHK(stage(bgS, gs, cnt * n, &dgS));
    VCHK(stage(bgVt, gvt, cnt * nn, &dgVt));
    VCHK(stage(bQ, qbar, cnt * mn, &dQ));
    if (have_fwd) {
        VCHK(stage(bU, u, cnt * mn, &dU));
        VCHK(stage(bS, s, cnt * n, &dS));
        VCHK(stage(bVt, vt, cnt * nn, &dVt));
    } else {  // recompute the forward pass instead of storing it
        VCHK(bU.alloc(sizeof(double) * cnt * mn));
        VCHK(bS.alloc(sizeof(double) * cnt * n));
        VCHK(bVt.alloc(sizeof(double) * cnt * nn));
        int rc = svd_device_uniform(h, count, m, n, dA, (double*)bU.p, (double*)bS.p, (double*)bVt.p);
        if (rc != LSVD_OK) return rc;
        dU = (const double*)bU.p;
        dS = (const double*)bS.p;
        dVt = (const double*)bVt.p;
    }
    double* dG = grad;
    if (host) {
        VCHK(bG.alloc(sizeof(double) * cnt * mn));
        dG = (double*)bG.p;
    }
    const size_t wsd = vjp_ws_doubles(m, n);
    VCHK(bW.alloc(sizeof(double) * wsd * cnt));
    VjpArgs va{count, m, n, dU, dS, dVt, dgU, dgS, dgVt, dQ, guard, dG, (double*)bW.p, wsd};
    VCHK(launch_svd_vjp(va, st));
    if (host) VCHK(hipMemcpyAsync(grad, dG, sizeof(double) * cnt * mn, hipMemcpyDeviceToHost, st));
    VCHK(hipStreamSynchronize(st));
#undef VCHK
    return LSVD_OK;
}

extern "C" {

int lsvd_svd_vjp_f64(lsvd_handle_t h, int count, int m, int n, const double* a, const double* u, const double* s,
                     const double* vt, const double* gu, const double* gs, const double* gvt, double guard,
                     double* grad) {
    return vjp_common(h, count, m, n, a, u, s, vt, gu, gs, gvt, nullptr, guard, grad);
}

int lsvd_polar_vjp_f64(lsvd_handle_t h, int count, int m, int n, const double* a, const double* qbar, double guard,
                       double* grad) {
    if (!qbar) return set_error(h, LSVD_E_ARG, "null qbar");
    return vjp_common(h, count, m, n, a, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, qbar, guard, grad);
}

}  // extern "C"
