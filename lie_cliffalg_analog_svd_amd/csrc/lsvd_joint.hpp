// lsvd_joint.hpp -- launch interface of the Phase-JADE kernels (joint.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/lie_svd_amd.h"

namespace lsvd {

struct JointRot {
    double c, s;  // rotor of the slot's pair
    int flag;     // 0 skipped (identity), 1 accepted, 2 rejected (the
                  // NaN pattern of the reference's rotate-and-roll-back)
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
// Layers [first, first + count) of a sweep (2 launches each), then the
// stop test when sweep_end.
int joint_layers(const JointArgs& a);
hipError_t joint_launch_layers(const JointArgs& a, hipStream_t st, int first, int count, bool sweep_end);
hipError_t joint_launch_diag(const JointArgs& a, double* diag, hipStream_t st);
// The schedule's (i, j) per (layer, slot) into device memory (layers x
// slots x 2 int32), by the device function the sweeps use.
hipError_t joint_launch_pairs(int n, int32_t* out, hipStream_t st);

// Largest number of kernel nodes in one captured graph.  A sweep of more
// layers is captured as several graphs replayed back to back.
constexpr int kMaxGraphNodes = 512;

// Splits `layers` layers of `nodes_per_layer` kernel nodes (plus one
// stop-test node at the end) into graphs of at most max_nodes nodes:
// writes the first layer and layer count of each graph, returns the number
// of graphs (at most max_chunks are written).
int sweep_graph_chunks(int layers, int nodes_per_layer, int max_nodes, int32_t* first, int32_t* count,
                       int max_chunks);

// One sweep (2 launches per layer + the stop test) captured as graphs of at
// most kMaxGraphNodes nodes each and relaunched on the handle's stream until
// every family has stopped or max_sweeps is reached; `fam` (host, a.F
// entries) receives the final per-family state.  Expects joint_launch_init
// to have run.
int joint_run_sweeps(lsvd_handle_t h, const JointArgs& a, int max_sweeps, JointFamily* fam);

}  // namespace lsvd
