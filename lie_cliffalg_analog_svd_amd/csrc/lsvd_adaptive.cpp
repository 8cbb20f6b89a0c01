// lsvd_adaptive.cpp -- host side of the adaptive triage and router
// (src/lie_svd_adaptive.rs): defaults, the batched triage call and the
// route decision on a triage.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "lsvd_adaptive.hpp"
#include "lsvd_api_internal.hpp"

using namespace lsvd;

namespace {

// temporaries come from the handle's workspace (lsvd_workspace.hpp)
using DevMem = lsvd::WsMem;

lsvd_adaptive_params_t defaults() {
    lsvd_adaptive_params_t p{};
    p.small_max_n = 512;
    p.coreflow_topo_max_n = 64;
    p.phaseflow_max_n = 64;
    p.nearly_diagonal_ratio = 1e-5;
    p.suspicious_threshold = 1.15;
    return p;
}

}  // namespace

extern "C" {

int lsvd_adaptive_params_default(lsvd_adaptive_params_t* out) {
    if (!out) return LSVD_E_ARG;
    *out = defaults();
    return LSVD_OK;
}

int lsvd_choose_route(const lsvd_triage_t* triage, const lsvd_adaptive_params_t* params) {
    if (!triage) return -1;
    return adaptive_route(*triage, params ? *params : defaults());
}

int lsvd_analyze_batched_f64(lsvd_handle_t h, int count, const int32_t* n, const int64_t* a_off, const double* a,
                             const lsvd_adaptive_params_t* params, lsvd_triage_t* out) {
    if (!h) return LSVD_E_ARG;
    if (count < 0) return set_error(h, LSVD_E_SHAPE, "negative count");
    if (count == 0) return LSVD_OK;
    if (!n || !a_off || !a || !out) return set_error(h, LSVD_E_ARG, "null pointer");
    std::vector<TriageItem> items((size_t)count);
    int64_t extent = 0;
    int max_n = 0;
    for (int b = 0; b < count; ++b) {
        if (n[b] < 0 || a_off[b] < 0) return set_error(h, LSVD_E_SHAPE, "adaptive SVD expects square matrices");
        if (n[b] > kTriageMaxN) return set_error(h, LSVD_E_LIMIT, "adaptive triage: n > 1024");
        items[b] = TriageItem{a_off[b], n[b], b};
        extent = std::max(extent, a_off[b] + (int64_t)n[b] * n[b]);
        max_n = std::max(max_n, (int)n[b]);
    }
    const lsvd_adaptive_params_t prm = params ? *params : defaults();
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(h, e, "hipSetDevice");
    hipStream_t st = h->stream;
#define ACHECK(expr)                                        \
    do {                                                    \
        hipError_t e_ = (expr);                             \
        if (e_ != hipSuccess) return hip_fail(h, e_, #expr); \
    } while (0)
    const bool host = h->ptr_mode == LSVD_PTR_HOST;
    lsvd::WsScope ws_scope(h);
    DevMem bitems, bpool, bout;
    ACHECK(bitems.alloc(sizeof(TriageItem) * (size_t)count));
    ACHECK(hipMemcpyAsync(bitems.p, items.data(), sizeof(TriageItem) * (size_t)count, hipMemcpyHostToDevice, st));
    const double* dpool = a;
    lsvd_triage_t* dout = out;
    if (host) {
        ACHECK(bpool.alloc(sizeof(double) * (size_t)extent));
        ACHECK(hipMemcpyAsync(bpool.p, a, sizeof(double) * (size_t)extent, hipMemcpyHostToDevice, st));
        ACHECK(bout.alloc(sizeof(lsvd_triage_t) * (size_t)count));
        dpool = (const double*)bpool.p;
        dout = (lsvd_triage_t*)bout.p;
    }
    ACHECK(launch_triage((const TriageItem*)bitems.p, count, max_n, dpool, prm, dout, st));
    if (host) ACHECK(hipMemcpyAsync(out, dout, sizeof(lsvd_triage_t) * (size_t)count, hipMemcpyDeviceToHost, st));
    ACHECK(hipStreamSynchronize(st));  // the temporaries are freed on return
#undef ACHECK
    return LSVD_OK;
}

}  // extern "C"
