// polish.hip -- batched square FP64 GEMMs on v_mfma_f64_16x16x4f64 for the
// digital-polish entry point: core = U0^T A V0 before the batched small SVD,
// U = U0 Ur and Vt = Vrt Vt0 after it.
//
// One 64-lane wavefront per 16 x 16 output tile; grid.x walks the tiles of
// the largest matrix, grid.y the matrices (tiles past a matrix's own size
// exit at once).  The operands come straight from global memory: a matrix of
// the batch is at most 1024^2 doubles and its tiles are re-read from L2.
#include <hip/hip_runtime.h>

#include "lsvd_polish.hpp"

namespace lsvd {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

template <bool TA, bool TB>
__global__ __launch_bounds__(64) void bgemm_kernel(const BgemmItem* items, const double* A, const double* B,
                                                   double* Cm) {
    const BgemmItem it = items[blockIdx.y];
    const int n = it.n;
    const int nt = (n + 15) >> 4;
    const int t = blockIdx.x;
    if (t >= nt * nt) return;
    const int I = t / nt, J = t % nt;
    const double* a = A + it.a_off;
    const double* b = B + it.b_off;
    double* c = Cm + it.c_off;
    const int lane = threadIdx.x;
    const int i = lane & 15, kk = lane >> 4;
    const int r = I * 16 + i, col = J * 16 + i;
    d4 acc = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < n; k0 += 4) {
        const int k = k0 + kk;
        // operand lanes: A[i][k] at lane i + 16 k', B[k][j] at lane j + 16 k'
        const double av = (r < n && k < n) ? (TA ? a[(int64_t)k * n + r] : a[(int64_t)r * n + k]) : 0.0;
        const double bv = (k < n && col < n) ? (TB ? b[(int64_t)col * n + k] : b[(int64_t)k * n + col]) : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
    }
    // accumulator element q of a lane: row 4 q + lane / 16, column lane % 16
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int row = I * 16 + 4 * q + kk, cc = J * 16 + i;
        if (row < n && cc < n) c[(int64_t)row * n + cc] = acc[q];
    }
}

}  // namespace

hipError_t launch_bgemm(const BgemmItem* items, int count, int max_n, const double* a, const double* b, double* c,
                        bool ta, bool tb, hipStream_t st) {
    if (count <= 0 || max_n <= 0) return hipSuccess;
    const int nt = (max_n + 15) / 16;
    for (int first = 0; first < count; first += 65535) {  // grid.y limit
        const dim3 grid(nt * nt, count - first < 65535 ? count - first : 65535);
        const BgemmItem* its = items + first;
        if (ta && !tb) hipLaunchKernelGGL((bgemm_kernel<true, false>), grid, dim3(64), 0, st, its, a, b, c);
        else if (!ta && tb) hipLaunchKernelGGL((bgemm_kernel<false, true>), grid, dim3(64), 0, st, its, a, b, c);
        else if (ta && tb) hipLaunchKernelGGL((bgemm_kernel<true, true>), grid, dim3(64), 0, st, its, a, b, c);
        else hipLaunchKernelGGL((bgemm_kernel<false, false>), grid, dim3(64), 0, st, its, a, b, c);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace lsvd
