// lsvd_types.hpp -- structures shared by host code and kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/lie_svd_amd.h"

namespace lsvd {

// One matrix of a ragged batch, as the kernels see it.  Offsets index the
// packed f64 pools (row-major matrices).
struct Item {
    int32_t m, n;     // caller's shape
    int32_t idx;      // caller's index (for info)
    int32_t mode;     // ItemMode
    int64_t a_off, u_off, s_off, vt_off;
};

enum ItemMode : int32_t {
    kModeSvd = 0,      // LieSvdSmall::solve / solve_rectangular
    kModeEigh = 1,     // eigh_jacobi_full on a symmetric square input
    kModeMicro = 2,    // LieSvdMicro (n <= 4 square)
    kModeMicroFallback = 3,  // LieSvdMicro escalated to LieSvdSmall (:90-92)
};

// Pool pointers of one batch execution.
struct Pools {
    const double* a;
    double* u;
    double* s;
    double* vt;
    lsvd_info_t* info;
    long long* timing;   // optional: kPhases clock64() deltas per matrix
};

// Phases recorded in Pools::timing (profiling aid).
enum Phase : int { kPhQR = 0, kPhNS = 1, kPhP = 2, kPhJacobi = 3, kPhU = 4, kPhRepair = 5, kPhOut = 6, kPhases = 8 };

// Solver constants of the reference (src/lie_svd_small.rs:67-69,
// src/lie_svd_micro.rs:15).
constexpr int kPolarMaxIters = 120;
constexpr double kPolarTol = 1e-13;
constexpr double kJacobiTol = 1e-13;
constexpr double kMicroTol = 1e-13;

// Step schedule of the Newton-Schulz iteration (LSVD_PREC_F64_SCALED and
// the low-precision phases).  The reference step X <- X (3I - X^T X) / 2
// is f(X) with f(x) = (3x - x^3) / 2 on the singular values; a scaled step
// applies f to s X (s = 1.5): Y = (3 / s^2) I - X^T X, X <- (s^3 / 2) X Y.
// Small singular values then grow by 1.5 s = 2.25 per step instead of 1.5,
// while those already near 1 stay in [f(s), 1] = [0.5625, 1].  Scaled steps
// run while ||X^T X - I||_F^2, compared two steps apart (the band of large
// values alternates), still falls by >= 0.25 (a small value arriving at
// the band removes ~1 from it); then plain steps finish with the
// reference's stopping rule, which is only tested on plain steps.  A plain
// phase whose step norm grows by > 1.3x per step while < 1e-3 (isolated
// small values still escaping) switches back.  Same limit, fewer steps.
struct NsSchedule {
    static constexpr double kS = 1.5;
    bool enabled, scaled;
    int flat;
    double g1, g2, dprev;
    __host__ __device__ void init(bool enable) {
        enabled = enable;
        scaled = enable;
        flat = 0;
        g1 = g2 = 1e300;
        dprev = -1.0;
    }
    // Y = alpha I - X^T X, X <- c X Y for the step about to run
    __host__ __device__ double alpha() const { return scaled ? 3.0 / (kS * kS) : 3.0; }
    __host__ __device__ double c() const { return scaled ? 0.5 * kS * kS * kS : 0.5; }
    // after a step: gi2 = ||X^T X - I||_F^2 of the iterate it started from,
    // delta = ||X_new - X||_F
    __host__ __device__ void after(double gi2, double delta) {
        if (scaled) {
            flat = (g2 < 1e299 && gi2 > g2 - 0.25) ? flat + 1 : 0;
            g2 = g1;
            g1 = gi2;
            if (flat >= 2 || gi2 < 0.05) {
                scaled = false;
                dprev = -1.0;
            }
        } else if (enabled) {
            if (dprev > 0.0 && delta < 1e-3 && delta > 1.3 * dprev) {
                scaled = true;
                flat = 0;
                g1 = g2 = 1e300;
            } else {
                dprev = delta;
            }
        }
    }
    // the reference's stopping rule applies to plain steps only
    __host__ __device__ bool plain() const { return !scaled; }
};

}  // namespace lsvd
