// lsvd_adaptive.hpp -- the adaptive router's decision rules, shared by the
// triage kernel (triage.hip) and the host entry points, and the kernel's
// launch interface.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lie_svd_amd.h"

namespace lsvd {

// is_balanced_degenerate_case (src/lie_svd_adaptive.rs:175-184).
__host__ __device__ inline bool adaptive_balanced_degenerate(const lsvd_triage_t& t, const lsvd_adaptive_params_t& p) {
    return t.suspicious_score >= p.suspicious_threshold + 0.25 && t.offdiag_ratio > 0.90 &&
           t.diagonal_dominance < 0.25 && t.row_cv >= 0.20 && t.col_cv >= 0.20 && t.row_cv < 0.45 &&
           t.col_cv < 0.45 && t.row_col_mismatch < 0.45;
}

// should_use_phaseflow (:121-150).
__host__ __device__ inline bool adaptive_phaseflow(const lsvd_triage_t& t, const lsvd_adaptive_params_t& p) {
    if (adaptive_balanced_degenerate(t, p)) return true;
    const bool causal = t.phase_causal_disbalance > 0.88 && t.offdiag_ratio > 0.20 && t.phase_mean_stress > 1.0;
    const bool locking = t.phase_mean_stress > 50.0 && t.phase_max_twist > 0.92 && t.row_cv < 0.42 &&
                         t.col_cv < 0.42 && t.row_col_mismatch < 0.30 && t.offdiag_ratio > 0.85 &&
                         t.diagonal_dominance < 0.35;
    const bool chiral = t.phase_chirality_balance > 0.30 && t.offdiag_ratio > 0.20 &&
                        t.phase_torsion_energy > 1.0 && t.diagonal_dominance < 0.5;
    return causal || locking || chiral;
}

// is_symmetric_topological_case + should_use_coreflow_topo (:152-173).
__host__ __device__ inline bool adaptive_coreflow_topo(const lsvd_triage_t& t, const lsvd_adaptive_params_t& p) {
    if (t.symmetry_ratio > 0.96 && t.offdiag_ratio > 0.35 && t.diagonal_dominance < 0.95) return true;
    if (adaptive_balanced_degenerate(t, p)) return true;
    const bool mass = t.row_mass_entropy < 0.92 || t.col_mass_entropy < 0.92;
    const bool balanced = t.row_col_mismatch < 0.22;
    const bool avoids = t.transpose_torsion_ratio < 0.62 || t.diagonal_entropy < 0.62;
    const bool not_rank1 = t.row_cv < 0.42 && t.col_cv < 0.42;
    return t.suspicious_score >= p.suspicious_threshold && mass && balanced && avoids && not_rank1;
}

// choose_route (:103-119); MICRO_MAX_N = 4 (src/lie_svd.rs:31).
__host__ __device__ inline int adaptive_route(const lsvd_triage_t& t, const lsvd_adaptive_params_t& p) {
    if (t.n <= 4) return LSVD_ROUTE_MICRO;
    if (t.offdiag_ratio <= p.nearly_diagonal_ratio) return LSVD_ROUTE_SMALL;
    if (t.n <= p.phaseflow_max_n && adaptive_phaseflow(t, p)) return LSVD_ROUTE_PHASEFLOW;
    if (t.n <= p.coreflow_topo_max_n && adaptive_coreflow_topo(t, p)) return LSVD_ROUTE_COREFLOW_TOPO;
    if (t.n <= p.small_max_n) return LSVD_ROUTE_SMALL;
    return LSVD_ROUTE_HYBRID;
}

struct TriageItem {
    int64_t off;  // doubles into the pool
    int n;
    int idx;      // position in the caller's list
};

constexpr int kTriageMaxN = 1024;

hipError_t launch_triage(const TriageItem* items, int count, int max_n, const double* pool,
                         const lsvd_adaptive_params_t& p, lsvd_triage_t* out, hipStream_t st);

}  // namespace lsvd
