// lsvd_joint.hpp -- launch interface of the Phase-JADE kernels (joint.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/lie_svd_amd.h"

namespace lsvd {

struct JointRot {
    double c, s;  // accepted rotor (identity when rejected or skipped)
    int flag;     // 1 when a rotation was accepted
};

// Per-family state (JointDiagonalizationTrace plus tolerances).
struct JointFamily {
    double pair_tol, tol;
    double initial_offdiag, final_offdiag;
    double before;  // off-diagonal norm at the start of the current sweep
    int sweeps, rotations, rejected, done;
};

struct JointArgs {
    int F, K, n, slots;
    const double* mats;  // F x K x n x n input (not modified)
    double* work;        // F x K x n x n working copy
    double* basis;       // F x n x n
    double* offrow;      // F x K x n off-diagonal row energies
    double* part;        // F x K x 2 (Frobenius^2, off-diagonal^2) partials
    JointRot* rot;       // F x slots, current layer
    JointFamily* fam;    // F
    double tol, min_pair_scale;
    int line_search_steps, max_sweeps;
    int rel_tol;  // 0: tolerances scale with max(||family||_F, 1) as the
                  // reference's; 1: with ||family||_F (eigh_jacobi_full's rule)
};

hipError_t joint_launch_init(const JointArgs& a, hipStream_t st);
hipError_t joint_launch_sweep(const JointArgs& a, hipStream_t st);
hipError_t joint_launch_diag(const JointArgs& a, double* diag, hipStream_t st);

// One sweep (2 launches per layer + the stop test) captured as a graph and
// relaunched on the handle's stream until every family has stopped or
// max_sweeps is reached; `fam` (host, a.F entries) receives the final
// per-family state.  Expects joint_launch_init to have run.
int joint_run_sweeps(lsvd_handle_t h, const JointArgs& a, int max_sweeps, JointFamily* fam);

}  // namespace lsvd
