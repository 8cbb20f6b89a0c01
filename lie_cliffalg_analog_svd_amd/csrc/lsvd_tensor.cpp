// lsvd_tensor.cpp -- LieSvdTensor::hosvd3 and rotate_tensor3
// (src/lie_svd_tensor.rs:31-62, 121-145) for a batch of same-shape tensors:
// the three mode Grams of every tensor go through LieSvdSmall::solve in one
// batched call of the small-SVD kernels.
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "lsvd_api_internal.hpp"
#include "lsvd_tensor.hpp"

using namespace lsvd;

namespace {

// temporaries come from the handle's workspace (lsvd_workspace.hpp)
using DMem = lsvd::WsMem;

int check_dims(lsvd_handle_t h, int count, int n1, int n2, int n3) {
    if (count < 0) return set_error(h, LSVD_E_SHAPE, "negative count");
    if (n1 <= 0 || n2 <= 0 || n3 <= 0) return set_error(h, LSVD_E_SHAPE, "empty tensor");
    if (n1 > kTensorMaxN || n2 > kTensorMaxN || n3 > kTensorMaxN)
        return set_error(h, LSVD_E_LIMIT, "tensor mode extent > " + std::to_string(kTensorMaxN));
    return LSVD_OK;
}

}  // namespace

#define TCHK(expr)                                          \
    do {                                                    \
        hipError_t e_ = (expr);                             \
        if (e_ != hipSuccess) return hip_fail(h, e_, #expr); \
    } while (0)

extern "C" {

int lsvd_hosvd3_f64(lsvd_handle_t h, int count, int n1, int n2, int n3, const double* x, double* core, double* u1,
                    double* u2, double* u3, lsvd_tensor_trace_t* trace) {
    if (!h) return LSVD_E_ARG;
    int rc = check_dims(h, count, n1, n2, n3);
    if (rc != LSVD_OK || count == 0) return rc;
    if (!x || !core || !u1 || !u2 || !u3) return set_error(h, LSVD_E_ARG, "null pointer");
    TCHK(hipSetDevice(h->device));
    hipStream_t st = h->stream;
    const bool host = h->ptr_mode == LSVD_PTR_HOST;
    const size_t N = (size_t)n1 * n2 * n3;
    const int dims[3] = {n1, n2, n3};
    const size_t per = (size_t)n1 * n1 + (size_t)n2 * n2 + (size_t)n3 * n3;  // one tensor's three Grams
    // the 3 count Grams: tensor t's mode m at t per + sum of earlier modes' n^2
    std::vector<int32_t> nv((size_t)3 * count);
    std::vector<int64_t> goff((size_t)3 * count), soff((size_t)3 * count);
    const size_t sper = (size_t)n1 + n2 + n3;
    for (int t = 0; t < count; ++t) {
        int64_t o = (int64_t)(t * per), so = (int64_t)(t * sper);
        for (int m = 0; m < 3; ++m) {
            nv[(size_t)3 * t + m] = dims[m];
            goff[(size_t)3 * t + m] = o;
            soff[(size_t)3 * t + m] = so;
            o += (int64_t)dims[m] * dims[m];
            so += dims[m];
        }
    }
    lsvd::WsScope ws_scope(h);
    DMem bx, bg, bu, bs, bvt, boff, bt1, bt2, bcore, bu1, bu2, bu3, bnorm;
    const double* dx = x;
    if (host) {
        TCHK(bx.alloc(sizeof(double) * N * count));
        TCHK(hipMemcpyAsync(bx.p, x, sizeof(double) * N * count, hipMemcpyHostToDevice, st));
        dx = (const double*)bx.p;
    }
    TCHK(bg.alloc(sizeof(double) * per * count));
    TCHK(bu.alloc(sizeof(double) * per * count));
    TCHK(bvt.alloc(sizeof(double) * per * count));
    TCHK(bs.alloc(sizeof(double) * sper * count));
    TCHK(boff.alloc(sizeof(int64_t) * 3 * count));
    TCHK(hipMemcpyAsync(boff.p, goff.data(), sizeof(int64_t) * 3 * count, hipMemcpyHostToDevice, st));
    const int64_t* doff = (const int64_t*)boff.p;
    TCHK(tensor_mode_grams(dx, count, n1, n2, n3, (double*)bg.p, doff, st));
    rc = svd_device_small(h, 3 * count, nv.data(), goff.data(), soff.data(), LSVD_PREC_F64, (const double*)bg.p,
                          (double*)bu.p, (double*)bs.p, (double*)bvt.p, nullptr);
    if (rc != LSVD_OK) return rc;
    // core = X x1 U1^T x2 U2^T x3 U3^T
    const double* du = (const double*)bu.p;
    double* dcore = core;
    if (host) {
        TCHK(bcore.alloc(sizeof(double) * N * count));
        dcore = (double*)bcore.p;
    }
    TCHK(bt1.alloc(sizeof(double) * N * count));
    TCHK(bt2.alloc(sizeof(double) * N * count));
    TCHK(tensor_mode_product(dx, count, n1, n2, n3, 0, n1, du, doff, 3, true, (double*)bt1.p, st));
    TCHK(tensor_mode_product((const double*)bt1.p, count, n1, n2, n3, 1, n2, du, doff, 3, true, (double*)bt2.p, st));
    TCHK(tensor_mode_product((const double*)bt2.p, count, n1, n2, n3, 2, n3, du, doff, 3, true, dcore, st));
    // U_m of every tensor out of the Gram-shaped pool
    double* outs[3] = {u1, u2, u3};
    DMem* stage[3] = {&bu1, &bu2, &bu3};
    size_t moff = 0;
    for (int m = 0; m < 3; ++m) {
        const size_t w = sizeof(double) * dims[m] * dims[m];
        double* dst = outs[m];
        if (host) {
            TCHK(stage[m]->alloc(w * count));
            dst = (double*)stage[m]->p;
        }
        TCHK(hipMemcpy2DAsync(dst, w, du + moff, sizeof(double) * per, w, count, hipMemcpyDeviceToDevice, st));
        moff += (size_t)dims[m] * dims[m];
    }
    if (trace) {
        TCHK(bnorm.alloc(sizeof(double) * 3 * count));
        double* dn = (double*)bnorm.p;  // [initial off | final off | ratio]
        TCHK(tensor_norms(dx, count, n1, n2, n3, dn, nullptr, st));
        TCHK(tensor_norms(dcore, count, n1, n2, n3, dn + count, dn + 2 * count, st));
        std::vector<double> hn((size_t)3 * count);
        TCHK(hipMemcpyAsync(hn.data(), dn, sizeof(double) * 3 * count, hipMemcpyDeviceToHost, st));
        TCHK(hipStreamSynchronize(st));
        for (int t = 0; t < count; ++t) {
            trace[t].initial_offdiag = hn[t];
            trace[t].final_offdiag = hn[(size_t)count + t];
            trace[t].superdiag_mass_ratio = hn[(size_t)2 * count + t];
        }
    }
    if (host) {
        TCHK(hipMemcpyAsync(core, dcore, sizeof(double) * N * count, hipMemcpyDeviceToHost, st));
        for (int m = 0; m < 3; ++m)
            TCHK(hipMemcpyAsync(outs[m], stage[m]->p, sizeof(double) * dims[m] * dims[m] * count,
                                hipMemcpyDeviceToHost, st));
    }
    TCHK(hipStreamSynchronize(st));
    return LSVD_OK;
}

int lsvd_tensor_rotate3_f64(lsvd_handle_t h, int count, int n1, int n2, int n3, int r1, int r2, int r3,
                            const double* x, const double* q1, const double* q2, const double* q3, double* out) {
    if (!h) return LSVD_E_ARG;
    int rc = check_dims(h, count, n1, n2, n3);
    if (rc == LSVD_OK) rc = check_dims(h, count, r1, r2, r3);
    if (rc != LSVD_OK || count == 0) return rc;
    if (!x || !q1 || !q2 || !q3 || !out) return set_error(h, LSVD_E_ARG, "null pointer");
    TCHK(hipSetDevice(h->device));
    hipStream_t st = h->stream;
    const bool host = h->ptr_mode == LSVD_PTR_HOST;
    const size_t nin = (size_t)n1 * n2 * n3, nout = (size_t)r1 * r2 * r3;
    const size_t qs[3] = {(size_t)r1 * n1, (size_t)r2 * n2, (size_t)r3 * n3};
    std::vector<int64_t> qoff((size_t)3 * count);
    for (int t = 0; t < count; ++t)
        for (int m = 0; m < 3; ++m) qoff[(size_t)3 * t + m] = (int64_t)(t * qs[m]);
    lsvd::WsScope ws_scope(h);
    DMem bx, bq[3], boff, bt1, bt2, bout;
    const double* dx = x;
    const double* dq[3] = {q1, q2, q3};
    double* dout = out;
    if (host) {
        TCHK(bx.alloc(sizeof(double) * nin * count));
        TCHK(hipMemcpyAsync(bx.p, x, sizeof(double) * nin * count, hipMemcpyHostToDevice, st));
        dx = (const double*)bx.p;
        for (int m = 0; m < 3; ++m) {
            TCHK(bq[m].alloc(sizeof(double) * qs[m] * count));
            TCHK(hipMemcpyAsync(bq[m].p, dq[m], sizeof(double) * qs[m] * count, hipMemcpyHostToDevice, st));
            dq[m] = (const double*)bq[m].p;
        }
        TCHK(bout.alloc(sizeof(double) * nout * count));
        dout = (double*)bout.p;
    }
    TCHK(boff.alloc(sizeof(int64_t) * 3 * count));
    TCHK(hipMemcpyAsync(boff.p, qoff.data(), sizeof(int64_t) * 3 * count, hipMemcpyHostToDevice, st));
    const int64_t* doff = (const int64_t*)boff.p;
    TCHK(bt1.alloc(sizeof(double) * (size_t)r1 * n2 * n3 * count));
    TCHK(bt2.alloc(sizeof(double) * (size_t)r1 * r2 * n3 * count));
    TCHK(tensor_mode_product(dx, count, n1, n2, n3, 0, r1, dq[0], doff, 3, false, (double*)bt1.p, st));
    TCHK(tensor_mode_product((const double*)bt1.p, count, r1, n2, n3, 1, r2, dq[1], doff, 3, false, (double*)bt2.p,
                             st));
    TCHK(tensor_mode_product((const double*)bt2.p, count, r1, r2, n3, 2, r3, dq[2], doff, 3, false, dout, st));
    if (host) TCHK(hipMemcpyAsync(out, dout, sizeof(double) * nout * count, hipMemcpyDeviceToHost, st));
    TCHK(hipStreamSynchronize(st));
    return LSVD_OK;
}

}  // extern "C"

#undef TCHK
