// lsvd_joint_svd.hpp -- launch interface of the two-sided family SVD
// kernels (joint_svd.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "lsvd_joint.hpp"

namespace lsvd {

struct JsvdRot {
    double cl, sl;  // left rotor (rows, and U's columns)
    double cr, sr;  // right rotor (columns, and V's columns)
    int flag;       // 1 when a rotation was accepted
};

struct JsvdArgs {
    int F, K, rows, cols, kk, slots;
    const double* mats;  // F x K x rows x cols input (not modified)
    double* work;        // F x K x rows x cols working copy
    double* ubasis;      // F x rows x rows
    double* vbasis;      // F x cols x cols
    double* rowoff;      // F x K x rows off-diagonal row energies
    double* coloff;      // F x K x cols off-diagonal column energies
    double* part;        // F x K x 2 (Frobenius^2, off-diagonal^2) partials
    JsvdRot* rot;        // F x slots, current layer
    JointFamily* fam;    // F
    double tol, min_pair_scale;
    int line_search_steps, max_sweeps;
};

hipError_t jsvd_launch_init(const JsvdArgs& a, hipStream_t st);
hipError_t jsvd_launch_sweep(const JsvdArgs& a, hipStream_t st);
hipError_t jsvd_launch_final(const JsvdArgs& a, double* sig, double* vt, hipStream_t st);

}  // namespace lsvd
