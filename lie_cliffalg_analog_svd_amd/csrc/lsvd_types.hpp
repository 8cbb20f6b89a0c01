// lsvd_types.hpp -- structures shared by host code and kernels.
#pragma once

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

}  // namespace lsvd
