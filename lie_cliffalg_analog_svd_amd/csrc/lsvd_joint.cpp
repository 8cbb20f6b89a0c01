// lsvd_joint.cpp -- host side of the Phase-JADE entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "lsvd_api_internal.hpp"
#include "lsvd_joint.hpp"
#include "lsvd_joint_svd.hpp"

using namespace lsvd;

namespace {

// temporaries come from the handle's workspace (lsvd_workspace.hpp)
struct DevBuf {
    void* p = nullptr;
    hipError_t ensure(size_t want) {
        lsvd::WsMem m;
        const hipError_t e = m.alloc(want);
        p = m.p;
        return e;
    }
};

}  // namespace

namespace lsvd {

int sweep_graph_chunks(int layers, int nodes_per_layer, int max_nodes, int32_t* first, int32_t* count,
                       int max_chunks) {
    if (layers <= 0) return 0;
    // layers per graph, leaving room for the stop-test node of the last one
    const int per = std::max(1, (max_nodes - 1) / std::max(nodes_per_layer, 1));
    int n = 0;
    for (int f = 0; f < layers; f += per, ++n) {
        if (n < max_chunks && first && count) {
            first[n] = f;
            count[n] = std::min(per, layers - f);
        }
    }
    return n;
}

// Captures one sweep as graphs of at most kMaxGraphNodes kernel nodes
// (launch(st, first, count, sweep_end) enqueues a layer range) and replays
// them until every family has stopped.  A single graph of a whole sweep
// reached 2047 nodes at n = 1024; under rocprofv3's kernel tracing such a
// launch ended in a host-side segmentation fault (graphs of up to 511 nodes
// replay under the tracer).
template <class Launch>
static int run_sweep_graphs(lsvd_handle_t h, int layers, const Launch& launch, int max_sweeps, const JointFamily* d_fam,
                            JointFamily* fam, int F, const char* what) {
    hipStream_t st = h->stream;
    const int nch = sweep_graph_chunks(layers, 2, kMaxGraphNodes, nullptr, nullptr, 0);
    std::vector<int32_t> first(nch), count(nch);
    sweep_graph_chunks(layers, 2, kMaxGraphNodes, first.data(), count.data(), nch);
    std::vector<hipGraph_t> graphs;
    std::vector<hipGraphExec_t> execs;
    auto cleanup = [&]() {
        for (hipGraphExec_t x : execs) (void)hipGraphExecDestroy(x);
        for (hipGraph_t g : graphs) (void)hipGraphDestroy(g);
    };
    for (int c = 0; c < nch; ++c) {
        hipGraph_t g = nullptr;
        hipGraphExec_t x = nullptr;
        hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) {
            cleanup();
            return hip_fail(h, e, "hipStreamBeginCapture");
        }
        hipError_t le = launch(st, first[c], count[c], c == nch - 1);
        e = hipStreamEndCapture(st, &g);
        if (e != hipSuccess || le != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            cleanup();
            return hip_fail(h, e != hipSuccess ? e : le, what);
        }
        graphs.push_back(g);
        e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            cleanup();
            return hip_fail(h, e, "hipGraphInstantiate");
        }
        execs.push_back(x);
    }
    int rc = LSVD_OK;
    for (int sweep = 0; sweep < max_sweeps; ++sweep) {
        hipError_t e = hipSuccess;
        for (int c = 0; c < nch && e == hipSuccess; ++c) e = hipGraphLaunch(execs[c], st);
        if (e == hipSuccess) e = hipMemcpyAsync(fam, d_fam, sizeof(JointFamily) * F, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            rc = hip_fail(h, e, what);
            break;
        }
        bool all = true;
        for (int f = 0; f < F; ++f) all = all && fam[f].done;
        if (all) break;
    }
    cleanup();
    return rc;
}

int joint_run_sweeps(lsvd_handle_t h, const JointArgs& a, int max_sweeps, JointFamily* fam) {
    auto launch = [&](hipStream_t st, int first, int count, bool end) {
        return joint_launch_layers(a, st, first, count, end);
    };
    return run_sweep_graphs(h, joint_layers(a), launch, max_sweeps, a.fam, fam, a.F, "joint sweep");
}

int jsvd_run_sweeps(lsvd_handle_t h, const JsvdArgs& a, int max_sweeps, JointFamily* fam) {
    auto launch = [&](hipStream_t st, int first, int count, bool end) {
        return jsvd_launch_layers(a, st, first, count, end);
    };
    return run_sweep_graphs(h, jsvd_layers(a), launch, max_sweeps, a.fam, fam, a.F, "joint svd sweep");
}

}  // namespace lsvd

extern "C" {

int lsvd_sweep_graph_chunks(int layers, int nodes_per_layer, int max_nodes, int32_t* first, int32_t* count,
                            int max_chunks) {
    return sweep_graph_chunks(layers, nodes_per_layer, max_nodes, first, count, max_chunks);
}

int lsvd_max_graph_nodes(void) { return kMaxGraphNodes; }

int lsvd_joint_layer_pairs(lsvd_handle_t h, int n, int32_t* out) {
    if (!h || !out) return LSVD_E_ARG;
    if (n < 0) return set_error(h, LSVD_E_SHAPE, "negative size");
    if (n > 4096) return set_error(h, LSVD_E_LIMIT, "n > 4096");
    const int layers = n % 2 == 0 ? n - 1 : n, slots = (n + 1) / 2;
    const size_t cnt = (size_t)std::max(layers, 0) * slots * 2;
    if (cnt == 0) return LSVD_OK;
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(h, e, "hipSetDevice");
    lsvd::WsScope ws_scope(h);
    DevBuf buf;
    if ((e = buf.ensure(sizeof(int32_t) * cnt)) != hipSuccess) return hip_fail(h, e, "hipMalloc");
    if ((e = joint_launch_pairs(n, (int32_t*)buf.p, h->stream)) != hipSuccess) return hip_fail(h, e, "pairs kernel");
    if ((e = hipMemcpyAsync(out, buf.p, sizeof(int32_t) * cnt, hipMemcpyDeviceToHost, h->stream)) != hipSuccess)
        return hip_fail(h, e, "hipMemcpyAsync");
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return hip_fail(h, e, "hipStreamSynchronize");
    return LSVD_OK;
}

// JointDiagonalizationParams::for_n (src/lie_svd_joint.rs:26-35).
int lsvd_joint_params_for_n(int n, lsvd_joint_params_t* out) {
    if (!out) return LSVD_E_ARG;
    out->max_sweeps = 12 + std::max(n, 1);
    out->tol = 1e-12;
    out->min_pair_scale = 1e-14;
    out->line_search_steps = 6;
    return LSVD_OK;
}

// Replaces LieSvdJoint::diagonalize_symmetric_with_params
// (src/lie_svd_joint.rs:71-158) for `families` independent families.
int lsvd_joint_diag_sym_f64(lsvd_handle_t h, int families, int k, int n, const double* mats,
                            const lsvd_joint_params_t* params, double* v, double* diag, lsvd_joint_trace_t* trace) {
    if (!h) return LSVD_E_ARG;
    if (families < 0 || n < 0) return set_error(h, LSVD_E_SHAPE, "negative size");
    if (k <= 0) return set_error(h, LSVD_E_SHAPE, "LieSvdJoint: empty matrix set");
    if (families == 0) return LSVD_OK;
    if (!mats || !v || !diag) return set_error(h, LSVD_E_ARG, "null pointer");
    if (n > 2048) return set_error(h, LSVD_E_LIMIT, "LieSvdJoint: n > 2048 (LDS row staging)");
    lsvd_joint_params_t prm;
    if (params) prm = *params;
    else lsvd_joint_params_for_n(n, &prm);
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(h, e, "hipSetDevice");
    hipStream_t st = h->stream;
    const size_t nn = (size_t)n * n;
    const size_t mats_n = (size_t)families * k * nn;
    const int slots = (n + (n & 1)) / 2;

    lsvd::WsScope ws_scope(h);
    DevBuf bmats, bwork, bbasis, boff, bpart, brot, bfam, bdiag;
#define JCHECK(expr)                                        \
    do {                                                    \
        hipError_t e_ = (expr);                             \
        if (e_ != hipSuccess) return hip_fail(h, e_, #expr); \
    } while (0)
    const bool host = h->ptr_mode == LSVD_PTR_HOST;
    const double* d_mats = mats;
    double* d_basis = v;
    double* d_diag = diag;
    if (host) {
        JCHECK(bmats.ensure(sizeof(double) * mats_n));
        JCHECK(hipMemcpyAsync(bmats.p, mats, sizeof(double) * mats_n, hipMemcpyHostToDevice, st));
        d_mats = (const double*)bmats.p;
        JCHECK(bbasis.ensure(sizeof(double) * families * nn));
        d_basis = (double*)bbasis.p;
        JCHECK(bdiag.ensure(sizeof(double) * (size_t)families * k * std::max(n, 1)));
        d_diag = (double*)bdiag.p;
    }
    JCHECK(bwork.ensure(sizeof(double) * std::max<size_t>(mats_n, 1)));
    JCHECK(boff.ensure(sizeof(double) * (size_t)families * k * std::max(n, 1)));
    JCHECK(bpart.ensure(sizeof(double) * (size_t)families * k * 2));
    JCHECK(brot.ensure(sizeof(JointRot) * (size_t)families * std::max(slots, 1)));
    JCHECK(bfam.ensure(sizeof(JointFamily) * families));

    JointArgs a{};
    a.F = families;
    a.K = k;
    a.n = n;
    a.slots = slots;
    a.mats = d_mats;
    a.work = (double*)bwork.p;
    a.basis = d_basis;
    a.offrow = (double*)boff.p;
    a.part = (double*)bpart.p;
    a.rot = (JointRot*)brot.p;
    a.fam = (JointFamily*)bfam.p;
    a.tol = prm.tol;
    a.min_pair_scale = prm.min_pair_scale;
    a.line_search_steps = prm.line_search_steps;
    a.max_sweeps = prm.max_sweeps;

    if (n == 0) {
        for (int f = 0; f < families && trace; ++f) trace[f] = lsvd_joint_trace_t{};
        return LSVD_OK;
    }
    JCHECK(joint_launch_init(a, st));
    std::vector<JointFamily> fam(families);
    if (n > 1 && prm.max_sweeps > 0) {
        // The families are independent: their sweeps run a chunk of
        // families at a time, sized so that a chunk's working copies and
        // bases stay in the Infinity Cache across the layers of a sweep
        // (every layer reads and writes all of them once).
        const size_t per_family = sizeof(double) * ((size_t)k + 1) * nn;
        const int chunk = h->jade_chunk_bytes == 0
                              ? families
                              : (int)std::max<size_t>(1, std::min<size_t>(families, h->jade_chunk_bytes / per_family));
        for (int f0 = 0; f0 < families; f0 += chunk) {
            JointArgs c = a;
            c.F = std::min(chunk, families - f0);
            c.mats = a.mats + (size_t)f0 * k * nn;
            c.work = a.work + (size_t)f0 * k * nn;
            c.basis = a.basis + (size_t)f0 * nn;
            c.offrow = a.offrow + (size_t)f0 * k * n;
            c.part = a.part + (size_t)f0 * k * 2;
            c.rot = a.rot + (size_t)f0 * slots;
            c.fam = a.fam + f0;
            int rc = joint_run_sweeps(h, c, prm.max_sweeps, fam.data() + f0);
            if (rc != LSVD_OK) return rc;
        }
    }
    JCHECK(joint_launch_diag(a, d_diag, st));
    JCHECK(hipMemcpyAsync(fam.data(), a.fam, sizeof(JointFamily) * families, hipMemcpyDeviceToHost, st));
    if (host) {
        JCHECK(hipMemcpyAsync(v, d_basis, sizeof(double) * families * nn, hipMemcpyDeviceToHost, st));
        JCHECK(hipMemcpyAsync(diag, d_diag, sizeof(double) * (size_t)families * k * n, hipMemcpyDeviceToHost, st));
    }
    JCHECK(hipStreamSynchronize(st));
#undef JCHECK
    if (trace) {
        for (int f = 0; f < families; ++f) {
            trace[f].initial_offdiag = n <= 1 ? 0.0 : fam[f].initial_offdiag;
            trace[f].final_offdiag = n <= 1 ? 0.0 : fam[f].final_offdiag;
            trace[f].sweeps = fam[f].sweeps;
            trace[f].rotations = fam[f].rotations;
            trace[f].rejected_rotations = fam[f].rejected;
        }
    }
    return LSVD_OK;
}

// Replaces LieSvdJoint::joint_svd_with_params (src/lie_svd_joint.rs:170-262)
// for `families` independent families of k rows x cols matrices.
int lsvd_joint_svd_f64(lsvd_handle_t h, int families, int k, int rows, int cols, const double* mats,
                       const lsvd_joint_params_t* params, double* u, double* sigmas, double* vt,
                       lsvd_joint_trace_t* trace) {
    if (!h) return LSVD_E_ARG;
    if (families < 0 || rows < 0 || cols < 0) return set_error(h, LSVD_E_SHAPE, "negative size");
    if (k <= 0) return set_error(h, LSVD_E_SHAPE, "LieSvdJoint: empty matrix set");
    if (rows == 0 || cols == 0) return set_error(h, LSVD_E_SHAPE, "LieSvdJoint: empty matrix");
    if (families == 0) return LSVD_OK;
    if (!mats || !u || !sigmas || !vt) return set_error(h, LSVD_E_ARG, "null pointer");
    if (cols > 4096) return set_error(h, LSVD_E_LIMIT, "LieSvdJoint::joint_svd: cols > 4096 (LDS column sums)");
    const int kk = std::min(rows, cols);
    lsvd_joint_params_t prm;
    if (params) prm = *params;
    else lsvd_joint_params_for_n(kk, &prm);
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(h, e, "hipSetDevice");
    hipStream_t st = h->stream;
    const size_t rc = (size_t)rows * cols, uu = (size_t)rows * rows, vv = (size_t)cols * cols;
    const size_t mats_n = (size_t)families * k * rc;
    const int slots = (kk + (kk & 1)) / 2;
    lsvd::WsScope ws_scope(h);
    DevBuf bmats, bwork, bu, bv, bsig, bvt, brow, bcol, bpart, brot, bfam;
#define JCHECK(expr)                                        \
    do {                                                    \
        hipError_t e_ = (expr);                             \
        if (e_ != hipSuccess) return hip_fail(h, e_, #expr); \
    } while (0)
    const bool host = h->ptr_mode == LSVD_PTR_HOST;
    const double* d_mats = mats;
    double* d_sig = sigmas;
    double* d_vt = vt;
    if (host) {
        JCHECK(bmats.ensure(sizeof(double) * mats_n));
        JCHECK(hipMemcpyAsync(bmats.p, mats, sizeof(double) * mats_n, hipMemcpyHostToDevice, st));
        d_mats = (const double*)bmats.p;
        JCHECK(bsig.ensure(sizeof(double) * (size_t)families * k * kk));
        d_sig = (double*)bsig.p;
        JCHECK(bvt.ensure(sizeof(double) * families * vv));
        d_vt = (double*)bvt.p;
    }
    JCHECK(bwork.ensure(sizeof(double) * mats_n));
    JCHECK(bu.ensure(sizeof(double) * families * uu));
    JCHECK(bv.ensure(sizeof(double) * families * vv));
    JCHECK(brow.ensure(sizeof(double) * (size_t)families * k * rows));
    const int rb = jsvd_blocks(jsvd_matrix_row_tasks(slots, rows, kk));
    JCHECK(bcol.ensure(sizeof(double) * (size_t)families * k * rb * cols));
    JCHECK(bpart.ensure(sizeof(double) * (size_t)families * k * 2));
    JCHECK(brot.ensure(sizeof(JsvdRot) * (size_t)families * std::max(slots, 1)));
    JCHECK(bfam.ensure(sizeof(JointFamily) * families));

    JsvdArgs a{};
    a.F = families;
    a.K = k;
    a.rows = rows;
    a.cols = cols;
    a.kk = kk;
    a.slots = slots;
    a.mats = d_mats;
    a.work = (double*)bwork.p;
    a.ubasis = (double*)bu.p;
    a.vbasis = (double*)bv.p;
    a.rowoff = (double*)brow.p;
    a.colpart = (double*)bcol.p;
    a.rb = rb;
    a.rbu = jsvd_blocks(jsvd_basis_row_tasks(rows));
    a.rbv = jsvd_blocks(jsvd_basis_row_tasks(cols));
    a.part = (double*)bpart.p;
    a.rot = (JsvdRot*)brot.p;
    a.fam = (JointFamily*)bfam.p;
    a.tol = prm.tol;
    a.min_pair_scale = prm.min_pair_scale;
    a.line_search_steps = prm.line_search_steps;
    a.max_sweeps = prm.max_sweeps;

    JCHECK(jsvd_launch_init(a, st));
    std::vector<JointFamily> fam(families);
    if (kk > 1 && prm.max_sweeps > 0) {
        int rc2 = jsvd_run_sweeps(h, a, prm.max_sweeps, fam.data());
        if (rc2 != LSVD_OK) return rc2;
    }
    JCHECK(jsvd_launch_final(a, d_sig, d_vt, st));
    JCHECK(hipMemcpyAsync(fam.data(), a.fam, sizeof(JointFamily) * families, hipMemcpyDeviceToHost, st));
    if (host) {
        JCHECK(hipMemcpyAsync(u, bu.p, sizeof(double) * families * uu, hipMemcpyDeviceToHost, st));
        JCHECK(hipMemcpyAsync(sigmas, d_sig, sizeof(double) * (size_t)families * k * kk, hipMemcpyDeviceToHost, st));
        JCHECK(hipMemcpyAsync(vt, d_vt, sizeof(double) * families * vv, hipMemcpyDeviceToHost, st));
    } else {
        JCHECK(hipMemcpyAsync(u, bu.p, sizeof(double) * families * uu, hipMemcpyDeviceToDevice, st));
    }
    JCHECK(hipStreamSynchronize(st));
#undef JCHECK
    if (trace) {
        for (int f = 0; f < families; ++f) {
            const bool trivial = kk <= 1;  // (:183-199): identity bases, zero trace
            trace[f].initial_offdiag = trivial ? 0.0 : fam[f].initial_offdiag;
            trace[f].final_offdiag = trivial ? 0.0 : fam[f].final_offdiag;
            trace[f].sweeps = trivial ? 0 : fam[f].sweeps;
            trace[f].rotations = trivial ? 0 : fam[f].rotations;
            trace[f].rejected_rotations = trivial ? 0 : fam[f].rejected;
        }
    }
    return LSVD_OK;
}

}  // extern "C"
