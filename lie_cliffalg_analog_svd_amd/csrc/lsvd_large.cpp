// lsvd_large.cpp -- LieSvdSmall::solve (src/lie_svd_small.rs:296-323) for one
// large square matrix with every step spread over the chip:
//   Q = newton_schulz_polar(A)            polar_run (MFMA GEMMs, all CUs)
//   P = sym(Q^T A)                        gemm_f64 + large_symmetrize
//   (V, eig) = eigh_jacobi_full(P)        block Jacobi (block_jacobi.hip):
//                                         Jacobi sweeps on block-pair
//                                         subproblems in LDS, their rotations
//                                         applied as MFMA tile products
//   sort, sigma = max(eig, 0), U = Q V    large_rank / large_permute + gemm_f64
// The workgroup-per-matrix kernels run each step on one CU; for a single
// matrix above a couple of hundred rows this route is the faster one.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "lsvd_api_internal.hpp"
#include "lsvd_large.hpp"
#include "lsvd_polar.hpp"

namespace lsvd {

namespace {

// temporaries come from the handle's workspace (lsvd_workspace.hpp)
using Mem = lsvd::WsMem;

}  // namespace

#define LCHK(expr)                                          \
    do {                                                    \
        hipError_t e_ = (expr);                             \
        if (e_ != hipSuccess) return hip_fail(h, e_, #expr); \
    } while (0)

namespace {

// eigh_jacobi_full (:212-285) of the symmetric n x n dP as a block Jacobi
// method over the whole chip (block_jacobi.hip), with the reference's stop
// test before every sweep: upper off-diagonal norm below JACOBI_TOL
// ||P||_F.  A sweep that no longer lowers the off-diagonal norm (the
// rounding floor) also ends the iteration.  Writes the unsorted diagonal
// dDiag (n) and the eigenvector columns dVc (n x n); synchronous.  A
// non-finite dP sets *finite = false and writes nothing (the caller hands
// such input to the workgroup kernels).
int run_block_jacobi(lsvd_handle_t h, const double* dP, int n, double* dDiag, double* dVc, int* sweeps_out,
                     int* rotations_out, bool* finite) {
    *finite = false;
    hipStream_t st = h->stream;
    const size_t nn = (size_t)n * n;
    const int bb = bj_tile_size(n);
    const int nbp = (n + bb - 1) / bb;
    const int ld = nbp * bb;
    const size_t ll = (size_t)ld * ld;
    const size_t wsz = (size_t)nbp * bb * bb;  // one layer's block rotations
    lsvd::WsScope ws_scope(h);
    Mem bApad, bVpad, bWrot, bPart, bRot;
    LCHK(bApad.alloc(sizeof(double) * ll));
    LCHK(bVpad.alloc(sizeof(double) * ll));
    LCHK(bWrot.alloc(sizeof(double) * wsz));
    LCHK(bPart.alloc(sizeof(double) * (kBjOffBlocks + 2)));
    LCHK(bRot.alloc(sizeof(int)));
    BlockJacobi bj{n, ld, 2 * nbp, bb, (double*)bApad.p, (double*)bVpad.p, (double*)bPart.p, (int*)bRot.p};
    double* wrot = (double*)bWrot.p;
    double* dsq = (double*)bPart.p + kBjOffBlocks;  // [0] ||P||_F^2, [1] off-diagonal^2
    LCHK(hipMemsetAsync(bRot.p, 0, sizeof(int), st));
    LCHK(bj_setup(bj, dP, st));
    LCHK(polar_sumsq(dP, nn, dsq, bj.partials, st));
    LCHK(bj_offnorm(bj, dsq + 1, st));
    double hq[2];
    LCHK(hipMemcpyAsync(hq, dsq, sizeof(hq), hipMemcpyDeviceToHost, st));
    LCHK(hipStreamSynchronize(st));
    if (!std::isfinite(hq[0]) || !std::isfinite(hq[1])) return LSVD_OK;
    *finite = true;
    const double jtol = 1e-13 * std::max(std::sqrt(hq[0]), 1e-300);
    double off = std::sqrt(hq[1]);
    const int max_sweeps = 60 + 20 * n;
    int sweeps = 0;
    while (sweeps < max_sweeps && off >= jtol) {
        for (int layer = 0; layer < bj.nb - 1; ++layer) {
            LCHK(bj_inner(bj, layer, wrot, st));
            LCHK(bj_update(bj, layer, wrot, st));
        }
        LCHK(bj_offnorm(bj, dsq + 1, st));
        LCHK(hipMemcpyAsync(hq, dsq, sizeof(hq), hipMemcpyDeviceToHost, st));
        LCHK(hipStreamSynchronize(st));
        ++sweeps;
        const double after = std::sqrt(hq[1]);
        const bool stalled = !(after < off * (1.0 - 1e-10));
        off = after;
        if (stalled) break;
    }
    int rotations = 0;
    LCHK(hipMemcpyAsync(&rotations, bRot.p, sizeof(int), hipMemcpyDeviceToHost, st));
    LCHK(bj_extract(bj, dDiag, dVc, st));
    LCHK(hipStreamSynchronize(st));
    *sweeps_out = sweeps;
    *rotations_out = rotations;
    return LSVD_OK;
}

int put_info(lsvd_handle_t h, lsvd_info_t* info, const lsvd_info_t& inf) {
    if (!info) return LSVD_OK;
    if (h->ptr_mode == LSVD_PTR_HOST) {
        *info = inf;
        return LSVD_OK;
    }
    LCHK(hipMemcpyAsync(info, &inf, sizeof(inf), hipMemcpyHostToDevice, h->stream));
    LCHK(hipStreamSynchronize(h->stream));
    return LSVD_OK;
}

}  // namespace

namespace {

// The route on device buffers for m >= n: dA m x n in; dU m x n, dS n,
// dVt n x n out.  *ok = false when the reference's repair branch applies
// (polar factor not trusted, or Q V not orthogonal to 1e-8 sqrt(n)).
int svd_large_tall(lsvd_handle_t h, const double* dA, int m, int n, double* dU, double* dS, double* dVt,
                   lsvd_info_t* inf, bool* ok) {
    *ok = false;
    hipStream_t st = h->stream;
    const size_t mn = (size_t)m * n, nn = (size_t)n * n;
    lsvd::WsScope ws_scope(h);
    Mem bQ, bP, bVs, bVc, bDiag, bPerm, bScal;
    LCHK(bQ.alloc(sizeof(double) * mn));
    lsvd_polar_info_t pinf{};
    int rc = polar_run(h, dA, m, n, LSVD_PREC_F64, (double*)bQ.p, &pinf);
    if (rc != LSVD_OK) return rc;
    if (!pinf.polar_ok) return LSVD_OK;  // the repair branch (:316-318)
    const double* dQ = (const double*)bQ.p;

    LCHK(bP.alloc(sizeof(double) * nn));
    double* dP = (double*)bP.p;
    LCHK(gemm_f64(true, n, n, m, dQ, dA, dP, 1.0, st));  // P = Q^T A
    LCHK(large_symmetrize(dP, n, st));

    LCHK(bDiag.alloc(sizeof(double) * n));
    LCHK(bVc.alloc(sizeof(double) * nn));
    double* dDiag = (double*)bDiag.p;
    double* dVc = (double*)bVc.p;
    int sweeps = 0, rotations = 0;
    bool finite = false;
    rc = run_block_jacobi(h, dP, n, dDiag, dVc, &sweeps, &rotations, &finite);
    if (rc != LSVD_OK || !finite) return rc;

    LCHK(bPerm.alloc(sizeof(int) * n));
    LCHK(bScal.alloc(sizeof(double) * 2));
    LCHK(bVs.alloc(sizeof(double) * nn));
    double* scal = (double*)bScal.p;  // [0] ||U^T U - I||_F^2, [1] non-finite flag (as int)
    LCHK(polar_zero(scal, 2, st));
    LCHK(large_rank(dDiag, n, (int*)bPerm.p, (int*)(scal + 1), st));
    LCHK(large_permute(dVc, dDiag, (const int*)bPerm.p, n, (double*)bVs.p, dVt, dS, true, st));
    LCHK(gemm_f64(false, m, n, n, dQ, (const double*)bVs.p, dU, 1.0, st));  // U = Q V
    LCHK(ns_orth(m, n, dU, scal, st));
    double hs[2];
    LCHK(hipMemcpyAsync(hs, scal, sizeof(hs), hipMemcpyDeviceToHost, st));
    LCHK(hipStreamSynchronize(st));
    int bad = 0;
    std::memcpy(&bad, &hs[1], sizeof(int));
    const double orth = std::sqrt(std::max(hs[0], 0.0));
    if (bad || !(orth <= 1e-8 * std::sqrt((double)n))) return LSVD_OK;  // the repair branch
    *inf = lsvd_info_t{};
    inf->polar_ok = 1;
    inf->ns_iters = pinf.f64_iters;
    inf->jacobi_sweeps = sweeps;
    inf->rotations = rotations;
    inf->path = 3;
    *ok = true;
    return LSVD_OK;
}

}  // namespace

int svd_large(lsvd_handle_t h, const double* a, int m, int n, double* u, double* s, double* vt, lsvd_info_t* info,
              bool* done) {
    *done = false;
    const int k = std::min(m, n);
    if (k < 2 || k > kLargeMaxN) return LSVD_OK;
    hipStream_t st = h->stream;
    const bool host = h->ptr_mode == LSVD_PTR_HOST;
    const bool wide = n > m;  // solved transposed (:348-352): SVD(A^T) = (V, sigma, U^T)
    const int tm = wide ? n : m, tn = k;  // the tall problem is tm x tn
    const size_t mn = (size_t)m * n, kk = (size_t)k * k;
    lsvd::WsScope ws_scope(h);
    Mem bA, bAt, bU, bS, bVt, bTu, bTvt;
    const double* dA = a;
    double *dU = u, *dS = s, *dVt = vt;
    if (host) {
        LCHK(bA.alloc(sizeof(double) * mn));
        LCHK(hipMemcpyAsync(bA.p, a, sizeof(double) * mn, hipMemcpyHostToDevice, st));
        dA = (const double*)bA.p;
        LCHK(bU.alloc(sizeof(double) * (size_t)m * k));
        LCHK(bS.alloc(sizeof(double) * k));
        LCHK(bVt.alloc(sizeof(double) * (size_t)k * n));
        dU = (double*)bU.p;
        dS = (double*)bS.p;
        dVt = (double*)bVt.p;
    }
    lsvd_info_t inf{};
    bool ok = false;
    int rc;
    if (!wide) {
        rc = svd_large_tall(h, dA, m, n, dU, dS, dVt, &inf, &ok);
    } else {
        // A^T (n x m) = U' S Vt'  =>  A = Vt'^T S U'^T: U = Vt'^T (m x m), Vt = U'^T (m x n)
        LCHK(bAt.alloc(sizeof(double) * mn));
        LCHK(bTu.alloc(sizeof(double) * (size_t)n * m));
        LCHK(bTvt.alloc(sizeof(double) * kk));
        LCHK(large_transpose(dA, m, n, (double*)bAt.p, st));
        rc = svd_large_tall(h, (const double*)bAt.p, tm, tn, (double*)bTu.p, dS, (double*)bTvt.p, &inf, &ok);
        if (rc == LSVD_OK && ok) {
            LCHK(large_transpose((const double*)bTvt.p, m, m, dU, st));
            LCHK(large_transpose((const double*)bTu.p, n, m, dVt, st));
        }
    }
    if (rc != LSVD_OK || !ok) return rc;
    if (host) {
        LCHK(hipMemcpyAsync(u, dU, sizeof(double) * (size_t)m * k, hipMemcpyDeviceToHost, st));
        LCHK(hipMemcpyAsync(s, dS, sizeof(double) * k, hipMemcpyDeviceToHost, st));
        LCHK(hipMemcpyAsync(vt, dVt, sizeof(double) * (size_t)k * n, hipMemcpyDeviceToHost, st));
    }
    LCHK(hipStreamSynchronize(st));
    rc = put_info(h, info, inf);
    if (rc == LSVD_OK) *done = true;
    return rc;
}

int eigh_large(lsvd_handle_t h, const double* p, int n, double* v, double* eig, lsvd_info_t* info, bool* done) {
    *done = false;
    if (n < 2 || n > kLargeMaxN) return LSVD_OK;
    hipStream_t st = h->stream;
    const bool host = h->ptr_mode == LSVD_PTR_HOST;
    const size_t nn = (size_t)n * n;
    lsvd::WsScope ws_scope(h);
    Mem bP, bVc, bDiag, bPerm, bFlag, bV, bE;
    const double* dP = p;
    double *dV = v, *dE = eig;
    if (host) {
        LCHK(bP.alloc(sizeof(double) * nn));
        LCHK(hipMemcpyAsync(bP.p, p, sizeof(double) * nn, hipMemcpyHostToDevice, st));
        dP = (const double*)bP.p;
        LCHK(bV.alloc(sizeof(double) * nn));
        LCHK(bE.alloc(sizeof(double) * n));
        dV = (double*)bV.p;
        dE = (double*)bE.p;
    }
    LCHK(bDiag.alloc(sizeof(double) * n));
    LCHK(bVc.alloc(sizeof(double) * nn));
    int sweeps = 0, rotations = 0;
    bool finite = false;
    int rc = run_block_jacobi(h, dP, n, (double*)bDiag.p, (double*)bVc.p, &sweeps, &rotations, &finite);
    if (rc != LSVD_OK || !finite) return rc;  // non-finite input: the workgroup kernels take it
    LCHK(bPerm.alloc(sizeof(int) * n));
    LCHK(bFlag.alloc(sizeof(double)));
    LCHK(polar_zero((double*)bFlag.p, 1, st));
    LCHK(large_rank((const double*)bDiag.p, n, (int*)bPerm.p, (int*)bFlag.p, st));
    LCHK(large_permute((const double*)bVc.p, (const double*)bDiag.p, (const int*)bPerm.p, n, dV, nullptr, dE, false,
                       st));
    int bad = 0;
    LCHK(hipMemcpyAsync(&bad, bFlag.p, sizeof(int), hipMemcpyDeviceToHost, st));
    LCHK(hipStreamSynchronize(st));
    if (bad) return LSVD_OK;  // non-finite input: the workgroup kernels report it as the reference does
    if (host) {
        LCHK(hipMemcpyAsync(v, dV, sizeof(double) * nn, hipMemcpyDeviceToHost, st));
        LCHK(hipMemcpyAsync(eig, dE, sizeof(double) * n, hipMemcpyDeviceToHost, st));
        LCHK(hipStreamSynchronize(st));
    }
    lsvd_info_t inf{};  // no polar step: polar_ok stays 0, as in the workgroup kernels' eigh mode
    inf.jacobi_sweeps = sweeps;
    inf.rotations = rotations;
    inf.path = 3;
    rc = put_info(h, info, inf);
    if (rc == LSVD_OK) *done = true;
    return rc;
}

#undef LCHK

}  // namespace lsvd
