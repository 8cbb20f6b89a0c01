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
    double* colpart;     // F x K x rb x cols column sums of squares (diagonal
                         // included), one partial per row block
    int rb;              // row blocks per matrix in the apply launch
    int rbu, rbv;        // row blocks of the U and V bases
    double* part;        // F x K x 2 (Frobenius^2, off-diagonal^2) partials
    JsvdRot* rot;        // F x slots, current layer
    JointFamily* fam;    // F
    double tol, min_pair_scale;
    int line_search_steps, max_sweeps;
};

// Apply launch geometry: one wavefront per row task, kJsvdTasks tasks per
// workgroup.  A matrix has the layer's slots (a row pair, or the odd index
// alone) and then its rows >= kk two at a time; a basis has its rows two
// at a time.
constexpr int kJsvdTasks = 4;
inline int jsvd_blocks(int tasks) { return (tasks + kJsvdTasks - 1) / kJsvdTasks; }
inline int jsvd_matrix_row_tasks(int slots, int rows, int kk) { return slots + (rows - kk + 1) / 2; }
inline int jsvd_basis_row_tasks(int n) { return (n + 1) / 2; }

hipError_t jsvd_launch_init(const JsvdArgs& a, hipStream_t st);
// Layers [first, first + count) of a sweep (2 launches each), then the
// stop test when sweep_end.
int jsvd_layers(const JsvdArgs& a);
hipError_t jsvd_launch_layers(const JsvdArgs& a, hipStream_t st, int first, int count, bool sweep_end);
hipError_t jsvd_launch_final(const JsvdArgs& a, double* sig, double* vt, hipStream_t st);

}  // namespace lsvd
