// lsvd_vjp.hpp -- launch interface of the SVD / polar adjoint kernel.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

namespace lsvd {

struct VjpArgs {
    int count, m, n;
    const double *u, *s, *vt;           // forward outputs (count x ...)
    const double *gu, *gs, *gvt;        // cotangents, each may be null (zero)
    const double* qbar;                 // polar mode: cotangent of Q = U Vt (null otherwise)
    double guard;
    double* grad;                       // count x m x n
    double* ws;                         // count x ws_stride doubles
    size_t ws_stride;
};

hipError_t launch_svd_vjp(const VjpArgs& a, hipStream_t st);
size_t vjp_ws_doubles(int m, int n);

// Thin SVD of `count` m x n (m >= n) matrices on device pointers (plain FP64
// LieSvdSmall path, no micro routing); synchronous on the handle's stream.
int svd_device_uniform(struct lsvd_handle_s* h, int count, int m, int n, const double* a, double* u, double* s,
                       double* vt);

}  // namespace lsvd
