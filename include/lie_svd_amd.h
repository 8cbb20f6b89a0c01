/*
 * lie_svd_amd.h -- C ABI of the MI355X (gfx950) dense SVD library.
 *
 * This is the drop-in boundary for the reference crate's dense SVD path
 * (skovnats/lie_cliffalg_analog_svd v0.44.0).  Each entry point names the
 * Rust function it replaces.  Conventions, fixed for every entry point:
 *
 *  - Matrices are row-major (C order) f64, the layout of ndarray's
 *    Array2<f64>.  Leading dimensions equal the column count.
 *  - Outputs follow the reference exactly: sigma is descending, Vt is
 *    returned (not V), tall inputs give thin U (m x k, k = min(m, n)),
 *    and LieSvdMicro's sign lives in U.
 *  - Pointer mode (lsvd_set_pointer_mode):
 *      LSVD_PTR_HOST   (default) all array arguments are host memory; the
 *                      call is synchronous, like the reference.
 *      LSVD_PTR_DEVICE array arguments are device memory; the call is
 *                      asynchronous on the handle's stream.
 *    Shape arrays (m[], n[], offsets) are always host memory.
 *  - The caller owns every buffer.  The library never frees caller memory.
 *  - Errors are return codes.  The reference asserts (panics) on shape
 *    misuse; a Rust wrapper maps LSVD_E_SHAPE back to a panic.  Numerical
 *    non-convergence is silent in the reference (fallbacks) and is reported
 *    here through lsvd_info_t, not as an error.
 *  - A handle is not thread-safe; separate handles are independent.
 */
#ifndef LIE_SVD_AMD_H
#define LIE_SVD_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lsvd_handle_s* lsvd_handle_t;
typedef struct lsvd_plan_s* lsvd_plan_t;

enum lsvd_status {
    LSVD_OK = 0,
    LSVD_E_SHAPE = 1,     /* shape misuse: the reference's assert! panics */
    LSVD_E_HIP = 2,       /* a HIP runtime call failed */
    LSVD_E_NONFINITE = 3, /* non-finite input where the path cannot proceed */
    LSVD_E_NOCONV = 4,    /* reserved: the path reports non-convergence in info */
    LSVD_E_ARG = 5,       /* bad argument (null pointer, unknown mode) */
    LSVD_E_NOMEM = 6,     /* device allocation failed */
    LSVD_E_LIMIT = 7      /* size beyond a documented maximum */
};

enum lsvd_pointer_mode { LSVD_PTR_HOST = 0, LSVD_PTR_DEVICE = 1 };

/* Arithmetic and step schedule of the Newton-Schulz polar iterations.
 * Every mode ends in FP64 with the reference's FP64 stopping rule.
 *   LSVD_PREC_F64        the reference iteration X <- X (3I - X^T X) / 2.
 *   LSVD_PREC_F64_SCALED the same limit in fewer FP64 steps: while small
 *                        singular values are still growing the step is
 *                        applied to 1.5 X (they grow 2.25x instead of
 *                        1.5x per step); plain steps finish.  No reference
 *                        counterpart.
 *   LSVD_PREC_F64_ONESIDED batch plans only: SVD items with max(m, n) <= 128
 *                        and 16 ceil(min(m, n) / 16) <= the handle's
 *                        os_max_np (default 128, environment LSVD_OS_MAX_NP at
 *                        handle creation, chosen by lsvd_autotune_batch) are
 *                        solved by Householder QR and one-sided (Hestenes)
 *                        Jacobi on R^T in FP64 (no Newton-Schulz; info.path =
 *                        4, ns_iters = 0); the rest of the plan runs as
 *                        LSVD_PREC_F64_SCALED.
 *   LSVD_PREC_F32 / BF16 low-precision MFMA iterations (FP32 accumulate,
 *                        scaled schedule), then FP64 refinement steps; large
 *                        polar only (lsvd_polar_f64), see there. */
enum lsvd_precision {
    LSVD_PREC_F64 = 0,
    LSVD_PREC_F32 = 1,
    LSVD_PREC_BF16 = 2,
    LSVD_PREC_F64_SCALED = 3,
    LSVD_PREC_F64_ONESIDED = 4
};

/* Per-matrix diagnostics (fills the reference's solver traces). */
typedef struct {
    int32_t polar_ok;       /* newton_schulz_polar's converged flag */
    int32_t ns_iters;       /* NS iterations (all precisions) */
    int32_t jacobi_sweeps;  /* Jacobi sweeps that applied rotations */
    int32_t rotations;      /* rotations applied */
    int32_t repaired;       /* left basis repaired from the right one */
    int32_t micro_fallback; /* LieSvdMicro escalated to LieSvdSmall */
    int32_t path;           /* 0 small, 1 micro, 2 micro->small fallback,
                               3 chip-wide route of a large single matrix,
                               4 QR + one-sided Jacobi (LSVD_PREC_F64_ONESIDED) */
    int32_t refine_iters;   /* FP64 refinement iterations (mixed precision) */
} lsvd_info_t;

/* JointDiagonalizationParams (src/lie_svd_joint.rs:18-35). */
typedef struct {
    int32_t max_sweeps;
    double tol;
    double min_pair_scale;
    int32_t line_search_steps;
} lsvd_joint_params_t;

/* JointDiagonalizationTrace (src/lie_svd_joint.rs:43-50). */
typedef struct {
    double initial_offdiag;
    double final_offdiag;
    int32_t sweeps;
    int32_t rotations;
    int32_t rejected_rotations;
} lsvd_joint_trace_t;

/* Polar diagnostics. */
typedef struct {
    int32_t polar_ok;      /* orth_err <= 1e-8 sqrt(n) (the reference's trust
                              test) and skew_err <= 1e-8 sqrt(n) */
    int32_t lowp_iters;    /* iterations in the chosen low precision */
    int32_t f64_iters;     /* FP64 iterations (refinement, or all for F64) */
    int32_t reserved;
    double orth_err;       /* ||Q^T Q - I||_F of the returned factor */
    double last_delta;     /* ||X_{k+1} - X_k||_F at the last iteration */
    double skew_err;       /* ||Q^T A - A^T Q||_F / ||A||_F: zero for A's polar
                              factor (Q^T A is then symmetric), so it sees the
                              error orthogonality cannot */
} lsvd_polar_info_t;

/* ---- handle ---------------------------------------------------------- */

/* Creates a handle on `device` with its own HIP stream and workspace. */
int lsvd_create(int device, lsvd_handle_t* out);
int lsvd_destroy(lsvd_handle_t h);
/* Uses a caller-owned hipStream_t (NULL restores the handle's own). */
int lsvd_set_stream(lsvd_handle_t h, void* hip_stream);
void* lsvd_get_stream(lsvd_handle_t h);
int lsvd_set_pointer_mode(lsvd_handle_t h, int mode);
/* No counterpart in the reference (its solves have one algorithm): the
 * precision mode of the single-matrix entry points below (LSVD_PREC_F64,
 * the reference iteration, by default; LSVD_PREC_F64_SCALED or
 * LSVD_PREC_F64_ONESIDED, else LSVD_E_ARG).  With LSVD_PREC_F64_ONESIDED a
 * single solve of max(m, n) <= 128 runs Householder QR + one-sided Jacobi
 * (as the batch route of that name).  The allocation-free entry point
 * always runs the reference iteration. */
int lsvd_set_solve_precision(lsvd_handle_t h, int precision);
int lsvd_synchronize(lsvd_handle_t h);
const char* lsvd_status_string(int status);
/* Text of the last error on this handle ("" if none). */
const char* lsvd_last_error(lsvd_handle_t h);
/* Library version string. */
const char* lsvd_version(void);

/* ---- LieSvdSmall ----------------------------------------------------- */

/* Replaces LieSvdSmall::solve (src/lie_svd_small.rs:296-323).
 * a: n x n (n <= 1024).  u: n x n, s: n, vt: n x n.  info may be NULL.
 * n <= 128 runs on one workgroup (LDS-resident kernels).  Larger n takes
 * the chip-wide route (info.path = 3): the polar iteration on the MFMA
 * GEMMs, P = sym(Q^T A), cyclic Jacobi over round-robin pair layers spread
 * over all CUs, the stable descending sort and U = Q V.  When the polar
 * factor is not trusted or Q V fails the orthogonality test, the solve
 * takes the reference's repair branch on the workgroup kernels (path 0).
 * The environment variable LSVD_LARGE_MIN_N, read when a handle is
 * created, moves the route's first n (values above 1024 disable it). */
int lsvd_small_solve_f64(lsvd_handle_t h, const double* a, int n, double* u, double* s, double* vt,
                         lsvd_info_t* info);

/* Replaces LieSvdSmall::solve_rectangular (src/lie_svd_small.rs:340-360).
 * a: m x n; k = min(m, n) <= 1024; u: m x k, s: k, vt: k x n.  k > 128
 * takes the chip-wide route (info.path = 3): the polar iteration on the
 * tall matrix itself (a wide one transposed), then as lsvd_small_solve_f64;
 * a rank-deficient input, whose polar factor is not trusted, goes to the
 * workgroup kernels' QR front end and repair (path 0). */
int lsvd_small_solve_rectangular_f64(lsvd_handle_t h, const double* a, int m, int n, double* u, double* s,
                                     double* vt, lsvd_info_t* info);

/* No counterpart in the reference: an allocation-free variant of
 * lsvd_small_solve_rectangular_f64 for callers that cannot allocate device
 * memory (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments and results (bit-identical: the same
 * on-chip kernel, at the default solve precision), but every pointer is a
 * device pointer (the handle must
 * be in LSVD_PTR_DEVICE mode, else LSVD_E_ARG), max(m, n) <= 128 (else
 * LSVD_E_LIMIT), and the call only enqueues work on the handle's stream.
 * The work queue, counter and staging slot live in a fixed 393,728-byte
 * device array that each GPU holds until the process ends; calls that
 * overlap on one GPU (other streams or host threads) race on it. */
int lsvd_small_solve_rectangular_f64_noalloc(lsvd_handle_t h, const double* a, int m, int n, double* u, double* s,
                                             double* vt, lsvd_info_t* info);

/* Replaces eigh_jacobi_full (src/lie_svd_small.rs:212-285): symmetric p
 * (n x n, n <= 1024) -> v (n x n, columns are eigenvectors), eig (n,
 * descending).  Above 128 (LSVD_LARGE_MIN_N) the Jacobi sweeps run as the
 * chip-wide block Jacobi method (info.path = 3). */
int lsvd_eigh_jacobi_f64(lsvd_handle_t h, const double* p, int n, double* v, double* eig, lsvd_info_t* info);

/* ---- LieSvdMicro ----------------------------------------------------- */

/* Replaces LieSvdMicro::solve (src/lie_svd_micro.rs:21-52), n <= 4. */
int lsvd_micro_solve_f64(lsvd_handle_t h, const double* a, int n, double* u, double* s, double* vt,
                         lsvd_info_t* info);

/* ---- metrics::compute (src/lib.rs:133-194) ---------------------------- */

/* SvdMetrics (src/lib.rs:136-142); sigma_max_rel / sigma_tail_rel are NaN
 * when no reference spectrum is given (the reference's None). */
typedef struct {
    double rel_recon;
    double orth_u;
    double orth_v;
    double sigma_max_rel;
    double sigma_tail_rel;
} lsvd_metrics_t;

/* Replaces metrics::compute (src/lib.rs:149-194) for a batch of
 * factorizations in the batch plan's pool layout (matrix b: a at a_off[b],
 * m x n; u at u_off[b], m x k; s at s_off[b], k; vt at vt_off[b], k x n;
 * k = min(m, n) <= 1024), computed on the device, one workgroup per matrix:
 * rel_recon = ||U diag(s) Vt - A||_F / max(||A||_F, 1e-300), orth_u =
 * ||U^T U - I||_F, orth_v = ||Vt Vt^T - I||_F (k x k identities: the
 * reference's n x n for its square factors), and with sigma_ref (same
 * layout as s, or NULL) the largest relative error of the descending-sorted
 * spectra over all i and over i >= floor(3 m / 4).  out[b] receives matrix
 * b's metrics.  Offsets are host arrays; pools and out follow the pointer
 * mode.  Synchronous. */
int lsvd_metrics_batched_f64(lsvd_handle_t h, int count, const int32_t* m, const int32_t* n, const int64_t* a_off,
                             const int64_t* u_off, const int64_t* s_off, const int64_t* vt_off, const double* a,
                             const double* u, const double* s, const double* vt, const double* sigma_ref,
                             lsvd_metrics_t* out);

/* ---- Newton-Schulz polar factor --------------------------------------- */

/* Replaces newton_schulz_polar (src/lie_svd_small.rs:158-183), which the
 * reference runs on square matrices, for tall or square m x n (m >= n):
 * q (m x n) gets the orthonormal polar factor.
 *   LSVD_PREC_F64  : the reference iteration in FP64 (FP64 MFMA products).
 *   LSVD_PREC_F64_SCALED : FP64 with the scaled step schedule (same limit).
 *   LSVD_PREC_F32  : FP32 MFMA products (scaled schedule) until the
 *   LSVD_PREC_BF16   iteration stalls, then FP64 refinement iterations
 *                    (BF16: BF16 MFMA, FP32 accumulate) to the reference's
 *                    FP64 stopping rule.
 * Low-precision products perturb the singular directions the FP64
 * refinement cannot restore: the result is orthogonal to FP64 precision
 * but is the polar factor of a matrix within about u_low ||A|| of A, so in
 * directions with sigma_i / sigma_max below the low precision's unit
 * roundoff u_low it is not A's.  info->orth_err is ||Q^T Q - I||_F and
 * info->skew_err the A-dependent residual ||Q^T A - A^T Q||_F / ||A||_F;
 * polar_ok needs both below 1e-8 sqrt(n), so such a factor reports 0.  The
 * Gram-Schmidt fallback of the reference is not run at this size: on
 * divergence polar_ok is 0 and q holds the last iterate. */
int lsvd_polar_f64(lsvd_handle_t h, const double* a, int m, int n, int precision, double* q,
                   lsvd_polar_info_t* info);

/* FP64 MFMA product C (m x n) = op(A) B, op(A) = A^T (A stored k x m) when
 * trans_a, else A (m x k); B is k x n.  Device pointer mode only.
 * (Diagnostic; no reference counterpart.) */
int lsvd_gemm_f64(lsvd_handle_t h, int trans_a, int m, int n, int k, const double* a, const double* b, double* c);

/* The same with FP32 (LSVD_PREC_F32) or BF16-rounded (LSVD_PREC_BF16)
 * operands and FP32 accumulation, FP32 data.  Device pointers only. */
int lsvd_gemm_lowp(lsvd_handle_t h, int precision, int trans_a, int m, int n, int k, const float* a,
                   const float* b, float* c);

/* ---- Adjoints (no reference counterpart: the reference has no
 *      derivatives) -------------------------------------------------------- */

/* Reverse-mode adjoint of the thin SVD (U, s, Vt) = solve_rectangular(A)
 * for `count` matrices of one shape m x n, m >= n (packed back to back).
 * u/s/vt: the forward outputs, or all three NULL to recompute the forward
 * pass (no storage of factors needed).  gu (m x n), gs (n), gvt (n x n):
 * cotangents of U, s and Vt; each may be NULL (zero).  grad (m x n) gets
 * dL/dA.  The 1/(s_i - s_j) factor of the singular-vector terms is guarded
 * as d/(d^2 + eps^2), eps = guard * s_max (guard 0 disables it; 1/(s_i +
 * s_j) and 1/s are treated the same way). */
int lsvd_svd_vjp_f64(lsvd_handle_t h, int count, int m, int n, const double* a, const double* u,
                     const double* s, const double* vt, const double* gu, const double* gs, const double* gvt,
                     double guard, double* grad);

/* Adjoint of the polar factor Q = U Vt of each m x n (m >= n) matrix:
 * qbar (m x n) is dL/dQ; grad gets dL/dA.  The forward SVD is recomputed. */
int lsvd_polar_vjp_f64(lsvd_handle_t h, int count, int m, int n, const double* a, const double* qbar,
                       double guard, double* grad);

/* ---- LieSvdJoint (Phase-JADE) ----------------------------------------- */

/* JointDiagonalizationParams::for_n (src/lie_svd_joint.rs:26-35). */
int lsvd_joint_params_for_n(int n, lsvd_joint_params_t* out);

/* Replaces LieSvdJoint::diagonalize_symmetric_with_params
 * (src/lie_svd_joint.rs:71-158), for `families` independent families at
 * once.  mats: families x k x n x n (family f, matrix q at
 * mats + (f*k + q)*n*n); v: families x n x n, the shared basis V with
 * diag(V^T M_q V) ~ diag_q; diag: families x k x n.  params NULL means
 * for_n(n).  trace (families entries, always host memory) may be NULL.
 * k == 0 is the reference's "empty matrix set" panic (LSVD_E_SHAPE).
 * n is limited to 2048 (LSVD_E_LIMIT beyond). */
int lsvd_joint_diag_sym_f64(lsvd_handle_t h, int families, int k, int n, const double* mats,
                            const lsvd_joint_params_t* params, double* v, double* diag,
                            lsvd_joint_trace_t* trace);

/* Replaces LieSvdJoint::joint_svd_with_params (src/lie_svd_joint.rs:
 * 170-262) for `families` independent families of k matrices of
 * rows x cols (mats: families x k x rows x cols, row-major).  Outputs per
 * family: u (rows x rows), sigmas (k x min(rows, cols), the |diagonal| of
 * U^T M V), vt (cols x cols, V^T) and the trace (JointSvdTrace,
 * :53-59, same fields as lsvd_joint_trace_t).  params NULL uses
 * lsvd_joint_params_for_n(min(rows, cols)).  The two-sided rotations pair
 * indices below min(rows, cols) only, as the reference does.  cols <= 4096
 * (LSVD_E_LIMIT beyond). */
int lsvd_joint_svd_f64(lsvd_handle_t h, int families, int k, int rows, int cols, const double* mats,
                       const lsvd_joint_params_t* params, double* u, double* sigmas, double* vt,
                       lsvd_joint_trace_t* trace);

/* Host memory only, no GPU work (no reference counterpart: the reference has
 * no graphs).  Phase-JADE and the joint SVD capture a sweep as hipGraphs of
 * at most lsvd_max_graph_nodes() kernel nodes each: `layers` layers of
 * `nodes_per_layer` nodes plus one stop-test node are split into graphs of
 * at most max_nodes nodes; first[c] / count[c] receive the layer range of
 * graph c (at most max_chunks entries written; both may be NULL).  Returns
 * the number of graphs. */
int lsvd_sweep_graph_chunks(int layers, int nodes_per_layer, int max_nodes, int32_t* first, int32_t* count,
                            int max_chunks);
int lsvd_max_graph_nodes(void);

/* Layer-pair map of the Phase-JADE and joint-SVD kernels: the pairs the
 * device's circle-method schedule (layer_pairs, src/lie_svd_joint.rs:
 * 516-539) yields for n indices, computed by a kernel with the same device
 * function the sweeps use.  out (host memory, layers x slots x 2 int32 with
 * layers = round_robin_layer_count(n), slots = ceil(n / 2)) receives (i, j),
 * i < j, per slot in the reference's order, or (-1, -1) for a slot paired
 * with the phantom index of odd n.  n <= 4096.  (Diagnostic.) */
int lsvd_joint_layer_pairs(lsvd_handle_t h, int n, int32_t* out);

/* Diagnostic, no reference counterpart: the device memory the handle keeps
 * for the temporaries of its entry points (the chip-wide route, the large
 * polar, Phase-JADE, the joint SVD, Phase-BSS, HO-SVD, polish, rotor export,
 * triage, vjp).  Those take their buffers from this workspace, which grows
 * to the largest call and is then reused, so repeated calls of one size
 * allocate nothing; it is freed by lsvd_destroy.  Either pointer may be
 * NULL. */
int lsvd_workspace_info(lsvd_handle_t h, size_t* bytes, int* chunks);

/* ---- Adaptive triage and router (src/lie_svd_adaptive.rs) ------------- */

/* AdaptiveRoute (src/lie_svd_adaptive.rs:8-15). */
enum lsvd_route {
    LSVD_ROUTE_MICRO = 0,
    LSVD_ROUTE_SMALL = 1,
    LSVD_ROUTE_PHASEFLOW = 2,
    LSVD_ROUTE_COREFLOW_TOPO = 3,
    LSVD_ROUTE_HYBRID = 4
};

/* LieSvdAdaptiveParams (src/lie_svd_adaptive.rs:59-78). */
typedef struct {
    int32_t small_max_n, coreflow_topo_max_n, phaseflow_max_n, reserved;
    double nearly_diagonal_ratio, suspicious_threshold;
} lsvd_adaptive_params_t;

/* AdaptiveTriage (src/lie_svd_adaptive.rs:18-50) and the route
 * choose_route picks from it (enum lsvd_route). */
typedef struct {
    int32_t n, route;
    double offdiag_ratio, diagonal_dominance, row_cv, col_cv, row_col_mismatch, symmetry_ratio,
        transpose_torsion_ratio, diagonal_entropy, row_mass_entropy, col_mass_entropy, phase_mean_stress,
        phase_max_twist, phase_causal_disbalance, phase_entropy_gap, phase_torsion_energy, phase_chirality_balance,
        phase_entropy, suspicious_score;
} lsvd_triage_t;

/* LieSvdAdaptiveParams::default (src/lie_svd_adaptive.rs:68-78). */
int lsvd_adaptive_params_default(lsvd_adaptive_params_t* out);

/* Replaces choose_route / analyze_matrix (src/lie_svd_adaptive.rs:103-119,
 * 186-274, with phase_signature src/lie_svd_phasehealth.rs:129-140 and the
 * Block4 dual balance src/lie_svd_block4.rs:210-262) for `count` square
 * matrices at once: matrix b is n[b] x n[b] at a + a_off[b] (host arrays;
 * 0 <= n[b] <= 1024, LSVD_E_LIMIT beyond).  out[b] receives the triage and
 * the route (params NULL: the defaults).  Sums run in another order than the
 * reference's loops and the device's log / atan2 / sin / cos differ from the
 * host's in the last bits, so a statistic sitting exactly on a threshold can
 * route differently. */
int lsvd_analyze_batched_f64(lsvd_handle_t h, int count, const int32_t* n, const int64_t* a_off, const double* a,
                             const lsvd_adaptive_params_t* params, lsvd_triage_t* out);

/* Host memory only.  choose_route's decision (src/lie_svd_adaptive.rs:
 * 103-184) on a triage (its route field is ignored); params NULL: the
 * defaults.  Returns the lsvd_route (-1 for a NULL triage). */
int lsvd_choose_route(const lsvd_triage_t* triage, const lsvd_adaptive_params_t* params);

/* ---- Tensor HO-SVD (batched consumer of LieSvdSmall) ------------------ */

/* TensorPhaseTrace (src/lie_svd_tensor.rs:11-17) without the shape. */
typedef struct {
    double initial_offdiag, final_offdiag, superdiag_mass_ratio;
} lsvd_tensor_trace_t;

/* Replaces LieSvdTensor::hosvd3 (src/lie_svd_tensor.rs:31-62) for `count`
 * tensors of one shape n1 x n2 x n3 (row-major, the last index fastest;
 * 1 <= n_i <= 64, LSVD_E_LIMIT beyond, LSVD_E_SHAPE for an empty one):
 * G_m = X_(m) X_(m)^T (the reference's summation order, no FMA), U_m from
 * LieSvdSmall::solve(G_m) -- all 3 count Grams in one batched call --, and
 * core = X x1 U1^T x2 U2^T x3 U3^T.  x, core: count n1 n2 n3 doubles; u1, u2,
 * u3: count n_m^2 each.  trace (host memory in both pointer modes, may be
 * NULL) receives count entries. */
int lsvd_hosvd3_f64(lsvd_handle_t h, int count, int n1, int n2, int n3, const double* x, double* core, double* u1,
                    double* u2, double* u3, lsvd_tensor_trace_t* trace);

/* ---- Phase-BSS (batched consumer of LieSvdSmall and LieSvdJoint) ------- */

/* PhaseBssTrace (src/lie_svd_bss.rs:36-44). */
typedef struct {
    int32_t channels, samples;
    double covariance_floor;
    double joint_initial_offdiag, joint_final_offdiag;
    int32_t joint_rotations;
    double mean_channel_coherence;
} lsvd_bss_trace_t;

/* Replaces LieSvdBss::separate (src/lie_svd_bss.rs:58-101) for `count`
 * observation blocks of one shape channels x samples (row-major, packed
 * back to back; 1 <= channels <= 64, LSVD_E_LIMIT beyond; samples >= 2,
 * else LSVD_E_SHAPE as the reference's assert): centered covariance, its
 * LieSvdSmall::solve (all blocks in one batched call), whitening
 * W[i][j] = u[j][i] / sqrt(max(sigma_i, eps max(max sigma, 1))), the
 * lagged covariances of the whitened channels for the lags 0 < lag <
 * samples (the covariance itself when none is usable), Phase-JADE of each
 * block's family (params NULL: for_n(channels)), then unmixing = V^T W
 * (channels^2), separated = V^T whitened (channels x samples) and the
 * channel coherence (channels) per block.  whitening / whitened (optional,
 * NULL to skip) receive W and the whitened channels.  lags: nlags host
 * ints (PhaseBssParams::for_channels uses 1, 2, 3, 5).  trace (host memory
 * in both pointer modes, may be NULL) receives count entries. */
int lsvd_bss_separate_f64(lsvd_handle_t h, int count, int channels, int samples, const double* x, int nlags,
                          const int* lags, double eps, const lsvd_joint_params_t* params, double* unmixing,
                          double* separated, double* coherence, double* whitening, double* whitened,
                          lsvd_bss_trace_t* trace);

/* Replaces rotate_tensor3 (src/lie_svd_tensor.rs:121-145), as used by
 * reconstruct_hosvd3 (:64-66): out[a][b][c] = sum q1[a][i] q2[b][j] q3[c][k]
 * x[i][j][k] for `count` tensors; q_m is r_m x n_m per tensor (count r_m n_m
 * doubles), out r1 x r2 x r3 per tensor.  Extents 1..64. */
int lsvd_tensor_rotate3_f64(lsvd_handle_t h, int count, int n1, int n2, int n3, int r1, int r2, int r3,
                            const double* x, const double* q1, const double* q2, const double* q3, double* out);

/* ---- Digital polish (batched consumer of LieSvdSmall) ---------------- */

/* Replaces the digital-polish step of the PhaseFlow, Analog and Hybrid
 * routes (src/lie_svd_phaseflow.rs:643-658, src/lie_svd_analog.rs:170-184,
 * src/lie_svd_hybrid.rs:108-116) for `count` square matrices at once:
 *   core = U0^T A V0 (V0 = Vt0^T), (Ur, sigma, Vrt) = LieSvdSmall::solve(core),
 *   U = U0 Ur, Vt = Vrt Vt0.
 * Matrix b is n[b] x n[b] (1 <= n[b] <= 1024); its A, U0, Vt0, U and Vt
 * blocks start at mat_off[b] of their pools and sigma at s_off[b] (host
 * arrays).  Every core takes the LieSvdSmall kernels, n <= 4 included, as
 * the reference calls LieSvdSmall::solve there.  precision: LSVD_PREC_F64
 * or LSVD_PREC_F64_SCALED for the cores' Newton-Schulz steps.  info (count
 * entries, the cores' diagnostics) may be NULL.  Synchronous in both
 * pointer modes. */
int lsvd_polish_batched_f64(lsvd_handle_t h, int count, const int32_t* n, const int64_t* mat_off,
                            const int64_t* s_off, const double* a, const double* u0, const double* vt0, int precision,
                            double* u, double* s, double* vt, lsvd_info_t* info);

/* ---- Rotor-stream export (src/lie_svd_compiler.rs) ---------------------- */

/* HardwareTarget (src/lie_svd_compiler.rs:12-25). */
enum lsvd_hw_target { LSVD_TARGET_MZI_MESH = 0, LSVD_TARGET_FPGA_ROTOR_MESH = 1 };

/* HardwarePhaseEvent::source and ::kind (src/lie_svd_compiler.rs:59-124,
 * 270-290): the event-producing routes and the rotor kinds they name. */
enum lsvd_event_source {
    LSVD_SRC_REAL_PHASEFLOW = 0,
    LSVD_SRC_COMPLEX_PHASEFLOW = 1,
    LSVD_SRC_ORTHOGONAL_GIVENS = 2
};
enum lsvd_event_kind {
    LSVD_KIND_GOLDEN_PRESPIN = 0,
    LSVD_KIND_CAUSAL_ANTISPIN = 1,
    LSVD_KIND_CROSS_PHASE_YINYANG = 2,
    LSVD_KIND_PHASE_CONJUGATE = 3,
    LSVD_KIND_BOTTLENECK = 4,
    LSVD_KIND_PHASE_JUMP = 5,
    LSVD_KIND_UNWRAP = 6,
    LSVD_KIND_DIRECTIONAL = 7,
    LSVD_KIND_SURGERY = 8,
    LSVD_KIND_COMPLEX_GOLDEN_PRESPIN = 9,
    LSVD_KIND_COMPLEX_DIAGONAL_PHASE = 10,
    LSVD_KIND_GIVENS_ELIMINATION = 11
};

/* HardwarePhaseEvent (src/lie_svd_compiler.rs:27-41).  flags bit 0:
 * energy_before present, bit 1: energy_after present (the Option fields). */
typedef struct {
    int32_t layer, i, j, flags;
    double phi_l, phi_r, theta, theta_l, theta_r;
    double energy_before, energy_after;
    int32_t source; /* enum lsvd_event_source */
    int32_t kind;   /* enum lsvd_event_kind */
} lsvd_phase_event_t;

/* Replaces the computation of HardwareSchedule::from_orthogonal_matrix
 * (src/lie_svd_compiler.rs:144-197) for `count` square matrices at once:
 * matrix b is d[b] x d[b] at v + v_off[b] (host arrays d and v_off;
 * transpose != 0 eliminates each matrix's transpose, e.g. V read from a Vt
 * pool).  Eliminations run in the reference's order (column j = 0..d-1,
 * rows i = j+1..d-1; entries with |r_ij| <= 1e-300 skipped) with its
 * arithmetic, so i, j, c, s and the signs equal the reference's bit for
 * bit; the reference's theta is atan2(s, c).  Outputs, packed in matrix
 * order (E_b = sum over earlier matrices of d (d - 1) / 2 event slots,
 * S_b = sum over earlier matrices of d sign slots):
 *   n_events[b]                         events of matrix b (event e's layer
 *                                       is e, as in the reference)
 *   ev_ij[2 (E_b + e)], [.. + 1]        i, j
 *   ev_cs[2 (E_b + e)], [.. + 1]        c, s
 *   ev_theta[E_b + e]                   atan2(s, c) by the device's atan2
 *                                       (NULL: not computed)
 *   signs[S_b + k]                      diagonal_signs: +-1, NaN for NaN
 * Slots past n_events[b] are left unwritten.  d[b] <= 1024 (LSVD_E_LIMIT). */
int lsvd_givens_schedule_f64(lsvd_handle_t h, int count, const int32_t* d, const int64_t* v_off, const double* v,
                             int transpose, int32_t* n_events, int32_t* ev_ij, double* ev_cs, double* ev_theta,
                             double* signs);

/* Host memory only.  The events from_orthogonal_matrix records for n
 * (i, j, c, s) tuples of lsvd_givens_schedule_f64 (src/lie_svd_compiler.rs:
 * 173-186): layer e, theta = theta_l = atan2(s, c) by the host's libm (the
 * function the reference calls), phi_l = phi_r = theta_r = 0, no energies. */
int lsvd_givens_events(int n, const int32_t* ev_ij, const double* ev_cs, lsvd_phase_event_t* out);

/* Host memory only.  HardwareSchedule::to_json_string
 * (src/lie_svd_compiler.rs:213-268), format "phase-schedule-v1": writes at
 * most cap - 1 bytes and a terminating NUL into buf (buf may be NULL when
 * cap is 0) and returns the full length without the NUL.  n_signs = 0
 * omits "diagonal_signs", as the reference does for an empty vector.
 * Numbers are printed as Rust's {:.17e} and the signs as Rust's {}.
 * Returns (size_t)-1 for an unknown target, source or kind. */
size_t lsvd_schedule_json(int target, int channels, int n_events, const lsvd_phase_event_t* events, int n_signs,
                          const double* signs, char* buf, size_t cap);

/* ---- Ragged batch (no reference counterpart: the reference does one
 *      matrix per call; this runs many LieSvd::solve / solve_rectangular
 *      calls, src/lie_svd.rs:49-51 + src/lie_svd_small.rs:340, at once) -- */

/* Builds a reusable plan for `count` matrices of shapes m[b] x n[b] (host
 * arrays).  Pools are packed: matrix b lives at a_off[b] in the input pool,
 * and its outputs at u_off[b] (m x k), s_off[b] (k) and vt_off[b] (k x n).
 * Square matrices with n <= 4 use the LieSvdMicro schedule, as the
 * reference's router does (src/lie_svd_adaptive.rs:93); the rest use
 * LieSvdSmall::solve_rectangular.  The plan sorts the work into buckets of
 * one persistent kernel launch each, picks tile shapes from the autotune
 * cache and spreads the buckets over up to four concurrent lanes.  precision is
 * LSVD_PREC_F64 (the reference iteration) or LSVD_PREC_F64_SCALED; the
 * low-precision modes are rejected here (LSVD_E_ARG): at the batch's
 * accuracy target (relative sigma error 1e-12 up to condition 1e12) a
 * low-precision iterate's error in the small-sigma directions is not
 * removed by FP64 refinement. */
int lsvd_batch_plan_create(lsvd_handle_t h, int count, const int32_t* m, const int32_t* n,
                           const int64_t* a_off, const int64_t* u_off, const int64_t* s_off,
                           const int64_t* vt_off, int precision, lsvd_plan_t* out);
/* Runs the plan.  Pool pointers follow the pointer mode; info (count
 * entries) may be NULL.  A plan with one lane replays a captured hipGraph;
 * a plan with concurrent lanes is enqueued on the handle's lane streams,
 * each with a hardware queue of its own, forked from and joined back to the
 * handle's stream with events.  Both forms may be captured into a caller's
 * graph (no memset nodes; counters are cleared by a kernel). */
int lsvd_batch_plan_execute(lsvd_plan_t p, const double* a_pool, double* u_pool, double* s_pool,
                            double* vt_pool, lsvd_info_t* info);
int lsvd_batch_plan_destroy(lsvd_plan_t p);
/* One-shot convenience: create + execute + destroy. */
int lsvd_svd_batched_f64(lsvd_handle_t h, int count, const int32_t* m, const int32_t* n, const int64_t* a_off,
                         const double* a_pool, const int64_t* u_off, double* u_pool, const int64_t* s_off,
                         double* s_pool, const int64_t* vt_off, double* vt_pool, int precision,
                         lsvd_info_t* info);

/* What lsvd_autotune_batch chose for the on-chip buckets (N, M <= 128) of
 * the last batch it tuned, and the plan times it measured (synthetic
 * U(-1, 1) inputs of the batch's shapes, median of 3 executions). */
typedef struct {
    int32_t multi_min_np;  /* padded sizes >= this share the multi-size launch (129: none) */
    int32_t max_lanes;     /* concurrent lanes (1..4) */
    int32_t wide_mask;     /* bit (np / 16 - 1): size-np bucket at twice the default threads */
    int32_t evaluations;   /* plans timed */
    double default_ms;     /* the handle's settings before tuning */
    double tuned_ms;       /* the chosen settings */
    int32_t os_max_np;     /* LSVD_PREC_F64_ONESIDED: padded sizes <= this take QR + one-sided Jacobi */
    int32_t lead_np;       /* padded size of the per-size bucket launched first (0: longest estimate first) */
} lsvd_tune_report_t;

/* Runs the autotuner for a batch's shapes and stores its choices in the
 * handle (used by later plans): for the on-chip buckets (N, M <= 128) a
 * coordinate search over (LSVD_PREC_F64_ONESIDED) the largest padded size
 * that takes QR + one-sided Jacobi, the split point of the multi-size launch
 * (64, 80, 96 or none), the lane count (1..4), the threads per workgroup
 * of each per-size bucket (default or twice) and which per-size bucket, if
 * any, is launched first, timing whole plans; for the buckets above 128 the
 * kernel variant of each bucket, timing that bucket. */
int lsvd_autotune_batch(lsvd_handle_t h, int count, const int32_t* m, const int32_t* n, int precision);
/* The plan-level part of lsvd_autotune_batch timed on the caller's own
 * device pools (the handle in LSVD_PTR_DEVICE mode, the layout of
 * lsvd_batch_plan_create), so that the search sees the data's real cost
 * mix (Newton-Schulz steps and Jacobi sweeps grow with the conditioning);
 * the outputs are overwritten by the timing runs. */
int lsvd_autotune_batch_pools(lsvd_handle_t h, int count, const int32_t* m, const int32_t* n, const int64_t* a_off,
                              const int64_t* u_off, const int64_t* s_off, const int64_t* vt_off, int precision,
                              const double* a_pool, double* u_pool, double* s_pool, double* vt_pool);
int lsvd_autotune_report(lsvd_handle_t h, lsvd_tune_report_t* out);

/* ---- Plan diagnostics (no reference counterpart: profiling aids) ------ */

/* One persistent launch of a plan.  variant -1 is the LieSvdMicro launch. */
typedef struct {
    int32_t first;      /* first queue entry of the bucket */
    int32_t count;      /* queue entries in the bucket */
    int32_t max_n;      /* largest min(m, n) in the bucket */
    int32_t variant;    /* kernel variant id (-1: micro) */
    int32_t grid;       /* workgroups launched */
    int32_t threads;    /* threads per workgroup */
    int64_t lds_bytes;  /* dynamic LDS per workgroup */
} lsvd_bucket_desc_t;

/* Writes up to `max` descriptors; *nbuckets gets the total.  order (may be
 * NULL) receives, for every queue entry, the caller's matrix index. */
int lsvd_batch_plan_describe(lsvd_plan_t p, lsvd_bucket_desc_t* out, int max, int* nbuckets, int32_t* order);

/* Runs the plan once without a graph, on the same concurrent lanes as
 * lsvd_batch_plan_execute, with HIP events recorded on each lane stream
 * around every launch,
 * and writes each launch's time in ms (same order as describe): the
 * duration of a bucket while the others run beside it.  Device pointer
 * mode only. */
int lsvd_batch_plan_time_buckets(lsvd_plan_t p, const double* a_pool, double* u_pool, double* s_pool,
                                 double* vt_pool, float* ms, int max);

/* Runs the plan once (no graph) with per-matrix phase clocks: `clocks` is
 * device memory of count x 8 int64, zeroed by the caller; entry [b][ph]
 * receives the clock64() cycles matrix b spent in phase ph (0 front end
 * (QR for N > 128), 1 Newton-Schulz, 2 P = Q^T A, 3 Jacobi, 4 U = Q V + check,
 * 5 repair, 6 output).  Device pointer mode only. */
int lsvd_batch_plan_phase_clocks(lsvd_plan_t p, const double* a_pool, double* u_pool, double* s_pool,
                                 double* vt_pool, int64_t* clocks);

#ifdef __cplusplus
}
#endif

#endif /* LIE_SVD_AMD_H */
