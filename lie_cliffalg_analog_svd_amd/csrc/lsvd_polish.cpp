// lsvd_polish.cpp -- host side of the batched digital-polish entry point:
// core = U0^T A V0, LieSvdSmall::solve(core), U = U0 Ur, Vt = Vrt Vt0, the
// finishing step of the PhaseFlow, Analog and Hybrid routes
// (src/lie_svd_phaseflow.rs:643-658, src/lie_svd_analog.rs:170-184,
// src/lie_svd_hybrid.rs:108-116), for a whole batch in one call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "lsvd_api_internal.hpp"
#include "lsvd_polish.hpp"

using namespace lsvd;

namespace {

// temporaries come from the handle's workspace (lsvd_workspace.hpp)
using DevMem = lsvd::WsMem;

}  // namespace

extern "C" {

int lsvd_polish_batched_f64(lsvd_handle_t h, int count, const int32_t* n, const int64_t* mat_off,
                            const int64_t* s_off, const double* a, const double* u0, const double* vt0, int precision,
                            double* u, double* s, double* vt, lsvd_info_t* info) {
    if (!h) return LSVD_E_ARG;
    if (count < 0) return set_error(h, LSVD_E_SHAPE, "negative count");
    if (count == 0) return LSVD_OK;
    if (!n || !mat_off || !s_off || !a || !u0 || !vt0 || !u || !s || !vt)
        return set_error(h, LSVD_E_ARG, "null pointer");
    if (precision != LSVD_PREC_F64 && precision != LSVD_PREC_F64_SCALED)
        return set_error(h, LSVD_E_ARG, "polish computes in FP64: precision must be LSVD_PREC_F64 or _F64_SCALED");
    int64_t mat_end = 0, s_end = 0;
    int max_n = 0;
    std::vector<BgemmItem> items((size_t)count);
    for (int b = 0; b < count; ++b) {
        if (n[b] < 1 || mat_off[b] < 0 || s_off[b] < 0)
            return set_error(h, LSVD_E_SHAPE, "LieSvdSmall::solve: matrix must be square N x N with N >= 1");
        if (n[b] > 1024) return set_error(h, LSVD_E_LIMIT, "polish: n > 1024");
        mat_end = std::max(mat_end, mat_off[b] + (int64_t)n[b] * n[b]);
        s_end = std::max(s_end, s_off[b] + (int64_t)n[b]);
        max_n = std::max(max_n, (int)n[b]);
        items[b] = BgemmItem{mat_off[b], mat_off[b], mat_off[b], n[b], 0};
    }
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(h, e, "hipSetDevice");
    hipStream_t st = h->stream;
#define PCHECK(expr)                                        \
    do {                                                    \
        hipError_t e_ = (expr);                             \
        if (e_ != hipSuccess) return hip_fail(h, e_, #expr); \
    } while (0)
    const bool host = h->ptr_mode == LSVD_PTR_HOST;
    const size_t mb = sizeof(double) * (size_t)mat_end, sb = sizeof(double) * (size_t)s_end;
    lsvd::WsScope ws_scope(h);
    DevMem bitems, bt, bcore, bur, bvrt, ba, bu0, bvt0, bu, bs, bvt, binfo;
    PCHECK(bitems.alloc(sizeof(BgemmItem) * (size_t)count));
    PCHECK(hipMemcpyAsync(bitems.p, items.data(), sizeof(BgemmItem) * (size_t)count, hipMemcpyHostToDevice, st));
    PCHECK(bt.alloc(mb));
    PCHECK(bcore.alloc(mb));
    PCHECK(bur.alloc(mb));
    PCHECK(bvrt.alloc(mb));
    const double *da = a, *du0 = u0, *dvt0 = vt0;
    double *du = u, *ds = s, *dvt = vt;
    lsvd_info_t* dinfo = info;
    if (host) {
        PCHECK(ba.alloc(mb));
        PCHECK(bu0.alloc(mb));
        PCHECK(bvt0.alloc(mb));
        PCHECK(bu.alloc(mb));
        PCHECK(bs.alloc(sb));
        PCHECK(bvt.alloc(mb));
        PCHECK(hipMemcpyAsync(ba.p, a, mb, hipMemcpyHostToDevice, st));
        PCHECK(hipMemcpyAsync(bu0.p, u0, mb, hipMemcpyHostToDevice, st));
        PCHECK(hipMemcpyAsync(bvt0.p, vt0, mb, hipMemcpyHostToDevice, st));
        // the caller's output pools go in too, so gaps between blocks come
        // back unchanged
        PCHECK(hipMemcpyAsync(bu.p, u, mb, hipMemcpyHostToDevice, st));
        PCHECK(hipMemcpyAsync(bs.p, s, sb, hipMemcpyHostToDevice, st));
        PCHECK(hipMemcpyAsync(bvt.p, vt, mb, hipMemcpyHostToDevice, st));
        da = (const double*)ba.p;
        du0 = (const double*)bu0.p;
        dvt0 = (const double*)bvt0.p;
        du = (double*)bu.p;
        ds = (double*)bs.p;
        dvt = (double*)bvt.p;
        if (info) {
            PCHECK(binfo.alloc(sizeof(lsvd_info_t) * (size_t)count));
            dinfo = (lsvd_info_t*)binfo.p;
        }
    }
    const BgemmItem* di = (const BgemmItem*)bitems.p;
    double* dt = (double*)bt.p;
    double* dcore = (double*)bcore.p;
    // core = (U0^T A) Vt0^T
    PCHECK(launch_bgemm(di, count, max_n, du0, da, dt, true, false, st));
    PCHECK(launch_bgemm(di, count, max_n, dt, dvt0, dcore, false, true, st));
    int rc = svd_device_small(h, count, n, mat_off, s_off, precision, dcore, (double*)bur.p, ds, (double*)bvrt.p,
                              dinfo);
    if (rc != LSVD_OK) return rc;
    // U = U0 Ur, Vt = Vrt Vt0
    PCHECK(launch_bgemm(di, count, max_n, du0, (const double*)bur.p, du, false, false, st));
    PCHECK(launch_bgemm(di, count, max_n, (const double*)bvrt.p, dvt0, dvt, false, false, st));
    if (host) {
        PCHECK(hipMemcpyAsync(u, du, mb, hipMemcpyDeviceToHost, st));
        PCHECK(hipMemcpyAsync(s, ds, sb, hipMemcpyDeviceToHost, st));
        PCHECK(hipMemcpyAsync(vt, dvt, mb, hipMemcpyDeviceToHost, st));
        if (info) PCHECK(hipMemcpyAsync(info, dinfo, sizeof(lsvd_info_t) * (size_t)count, hipMemcpyDeviceToHost, st));
    }
    PCHECK(hipStreamSynchronize(st));  // the temporaries are freed on return
#undef PCHECK
    return LSVD_OK;
}

}  // extern "C"
