// vjp.hip -- reverse-mode adjoint of the thin SVD A = U diag(s) Vt
// (A m x n, m >= n, U m x n, Vt n x n) and of the polar factor Q = U Vt.
// No reference counterpart: the reference has no derivatives.
//
// With cotangents gU (m x n), gs (n), gVt (n x n) and V = Vt^T:
//   J_U = U^T gU,  J_V = V^T gV = Vt gVt^T,  Om_U = skew(J_U), Om_V = skew(J_V)
//   M_ij = (Om_U + Om_V)_ij / (s_j - s_i) + (Om_U - Om_V)_ij / (s_j + s_i)   (i != j)
//   M_ii = gs_i
//   dA   = U (M - J_U diag(1/s)) Vt + gU diag(1/s) Vt
// which is the textbook F_ij = 1/(s_j^2 - s_i^2) form split into its
// difference and sum parts.  Only 1/(s_j - s_i) can blow up at a
// near-degenerate gap; it is guarded as d / (d^2 + eps^2) with
// eps = guard * s_max (and 1/s as s / (s^2 + eps^2)).  For losses that are
// invariant to the singular-vector gauge (e.g. <G, U Vt>) Om_U + Om_V = 0
// exactly and the guard never enters; the polar mode drops that term.
//
// One workgroup per matrix; the n x n and m x n intermediates live in a
// per-matrix global workspace (L2 resident) and the products use the
// in-workgroup GEMM.
#include <hip/hip_runtime.h>
#include <math.h>

#include "lsvd_device.hpp"
#include "lsvd_vjp.hpp"

namespace lsvd {

struct ZeroView {
    __device__ __forceinline__ double operator()(int, int) const { return 0.0; }
};

template <int NT>
__global__ __launch_bounds__(NT) void svd_vjp_kernel(VjpArgs a) {
    const int b = blockIdx.x;
    const int m = a.m, n = a.n;
    const size_t mn = (size_t)m * n, nn = (size_t)n * n;
    const double* U = a.u + b * mn;
    const double* s = a.s + (size_t)b * n;
    const double* Vt = a.vt + b * nn;
    double* ws = a.ws + (size_t)b * a.ws_stride;
    double* JU = ws;            // n x n
    double* JV = JU + nn;       // n x n
    double* Nm = JV + nn;       // n x n
    double* T = Nm + nn;        // m x n
    double* gUw = T + mn;       // m x n (polar mode: Qbar V)
    double* gVw = gUw + mn;     // n x n (polar mode: U^T Qbar)
    __shared__ double red[NT / 64 + 2];
    __shared__ double smax_s;

    const double* gU = a.gu ? a.gu + b * mn : nullptr;
    const double* gVt = a.gvt ? a.gvt + b * nn : nullptr;
    const double* gs = a.gs ? a.gs + (size_t)b * n : nullptr;
    if (a.qbar) {
        // polar Q = U Vt: gU = Qbar V (m x n), gVt = U^T Qbar (n x n), gs = 0
        const double* Qb = a.qbar + b * mn;
        wg_gemm<NT>(m, n, n, View{Qb, n, 0}, View{Vt, n, 1}, [&](int i, int j, double acc) { gUw[(size_t)i * n + j] = acc; });
        wg_gemm<NT>(n, n, m, View{U, n, 1}, View{Qb, n, 0}, [&](int i, int j, double acc) { gVw[(size_t)i * n + j] = acc; });
        __syncthreads();
        gU = gUw;
        gVt = gVw;
        gs = nullptr;
    }
    // J_U = U^T gU, J_V = Vt gVt^T
    if (gU)
        wg_gemm<NT>(n, n, m, View{U, n, 1}, View{gU, n, 0}, [&](int i, int j, double acc) { JU[(size_t)i * n + j] = acc; });
    else
        for (int e = threadIdx.x; e < n * n; e += NT) JU[e] = 0.0;
    if (gVt)
        wg_gemm<NT>(n, n, n, View{Vt, n, 0}, View{gVt, n, 1}, [&](int i, int j, double acc) { JV[(size_t)i * n + j] = acc; });
    else
        for (int e = threadIdx.x; e < n * n; e += NT) JV[e] = 0.0;
    double mx = 0.0;
    for (int i = threadIdx.x; i < n; i += NT) mx = fmax(mx, fabs(s[i]));
    mx = block_max<NT>(mx, red);
    if (threadIdx.x == 0) smax_s = mx;
    __syncthreads();
    const double eps = a.guard * fmax(smax_s, 1e-300);
    const double eps2 = eps * eps;
    // N = M - J_U diag(1/s)
    for (int e = threadIdx.x; e < n * n; e += NT) {
        const int i = e / n, j = e % n;
        const double si = s[i], sj = s[j];
        const double inv_sj = sj / (sj * sj + eps2);
        double mij;
        if (i == j) {
            mij = gs ? gs[i] : 0.0;
        } else {
            const double ou = 0.5 * (JU[(size_t)i * n + j] - JU[(size_t)j * n + i]);
            const double ov = 0.5 * (JV[(size_t)i * n + j] - JV[(size_t)j * n + i]);
            const double d = sj - si, su = sj + si;
            // for the polar factor Om_U + Om_V vanishes identically; its
            // rounding residue is left out instead of being divided by a gap
            const double dif = a.qbar ? 0.0 : (ou + ov) * d / (d * d + eps2);
            mij = dif + (ou - ov) * su / (su * su + eps2);
        }
        Nm[(size_t)i * n + j] = mij - JU[(size_t)i * n + j] * inv_sj;
    }
    __syncthreads();
    // T = U N + gU diag(1/s)
    wg_gemm<NT>(m, n, n, View{U, n, 0}, View{Nm, n, 0}, [&](int i, int j, double acc) {
        const double sj = s[j];
        T[(size_t)i * n + j] = acc + (gU ? gU[(size_t)i * n + j] * (sj / (sj * sj + eps2)) : 0.0);
    });
    __syncthreads();
    // dA = T Vt
    double* G = a.grad + b * mn;
    wg_gemm<NT>(m, n, n, View{T, n, 0}, View{Vt, n, 0}, [&](int i, int j, double acc) { G[(size_t)i * n + j] = acc; });
}

hipError_t launch_svd_vjp(const VjpArgs& a, hipStream_t st) {
    if (a.count <= 0) return hipSuccess;
    hipLaunchKernelGGL(svd_vjp_kernel<256>, dim3(a.count), dim3(256), 0, st, a);
    return hipGetLastError();
}

size_t vjp_ws_doubles(int m, int n) {
    const size_t mn = (size_t)m * n, nn = (size_t)n * n;
    return ((4 * nn + 2 * mn) + 31) / 32 * 32;
}

}  // namespace lsvd
