// large.hip -- the element-wise steps of LieSvdSmall::solve for one large
// square matrix run across the whole chip (src/lie_svd_small.rs:296-323):
// symmetrising P = Q^T A (:304-308), the descending stable order of the
// Jacobi diagonal with the column permutation of V (eigh_jacobi_full's sort,
// :268-284), V^T, and sigma = max(eig, 0) (:312-315).  The products are the
// polar GEMMs; the Jacobi sweeps are the block Jacobi of block_jacobi.hip.
#include <hip/hip_runtime.h>
#include <math.h>

#include "lsvd_large.hpp"

namespace lsvd {
namespace {

constexpr int kT = 32;     // tile edge
constexpr int kRows = 8;   // thread rows per tile (blockDim = 32 x 8)

// P <- (P + P^T) * 0.5 for tile pair (bi, bj), bj >= bi: both tiles are
// read into LDS before either is written, and no other workgroup touches
// them.  The sum p_ij + p_ji is formed once per entry pair, so the result
// is exactly symmetric and the diagonal is unchanged.
__global__ __launch_bounds__(kT * kRows) void large_sym_kernel(double* P, int n) {
    const int bi = blockIdx.y, bj = blockIdx.x;
    if (bj < bi) return;
    __shared__ double ta[kT][kT + 1], tb[kT][kT + 1];
    const int tx = threadIdx.x;
    for (int r = threadIdx.y; r < kT; r += kRows) {
        const int i = bi * kT + r, j = bj * kT + tx;
        ta[r][tx] = (i < n && j < n) ? P[(size_t)i * n + j] : 0.0;
        const int i2 = bj * kT + r, j2 = bi * kT + tx;
        tb[r][tx] = (i2 < n && j2 < n) ? P[(size_t)i2 * n + j2] : 0.0;
    }
    __syncthreads();
    for (int r = threadIdx.y; r < kT; r += kRows) {
        const int i = bi * kT + r, j = bj * kT + tx;
        if (i < n && j < n) P[(size_t)i * n + j] = (ta[r][tx] + tb[tx][r]) * 0.5;
        const int i2 = bj * kT + r, j2 = bi * kT + tx;
        if (bi != bj && i2 < n && j2 < n) P[(size_t)i2 * n + j2] = (tb[r][tx] + ta[tx][r]) * 0.5;
    }
}

// Stable descending order of d (one workgroup, n <= kLargeMaxN): entry i
// goes to rank #{j : d_j > d_i} + #{j < i : d_j == d_i}, which is the
// position a stable sort by "d_j compared with d_i" gives it.  A non-finite
// entry sets *bad (its rank would collide), and the caller does not use
// the permutation then.
__global__ __launch_bounds__(1024) void large_rank_kernel(const double* d, int n, int* perm, int* bad) {
    __shared__ double s[kLargeMaxN];
    __shared__ int flag;
    if (threadIdx.x == 0) flag = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s[i] = d[i];
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double di = s[i];
        if (!isfinite(di)) flag = 1;
        int r = 0;
        for (int j = 0; j < n; ++j) r += (s[j] > di) || (s[j] == di && j < i);
        if (r < n) perm[r] = i;
    }
    __syncthreads();
    if (threadIdx.x == 0) *bad = flag;
}

// Vs[:, k] = V[:, perm[k]], Vt = Vs^T (through an LDS tile; skipped when
// Vt is null), and sigma[k] = d[perm[k]], clamped to max(., 0) for the SVD.
// Tile (bi, bj): rows bi of V, sorted columns bj.
__global__ __launch_bounds__(kT * kRows) void large_permute_kernel(const double* V, const double* d, const int* perm,
                                                                  int n, double* Vs, double* Vt, double* sigma,
                                                                  int clamp) {
    __shared__ double t[kT][kT + 1];
    const int bi = blockIdx.y, bj = blockIdx.x, tx = threadIdx.x;
    const int col = bj * kT + tx;
    const int src = col < n ? perm[col] : 0;
    for (int r = threadIdx.y; r < kT; r += kRows) {
        const int i = bi * kT + r;
        double v = 0.0;
        if (i < n && col < n) {
            v = V[(size_t)i * n + src];
            Vs[(size_t)i * n + col] = v;
        }
        t[r][tx] = v;
    }
    __syncthreads();
    if (Vt) {
        for (int r = threadIdx.y; r < kT; r += kRows) {
            const int k = bj * kT + r, i = bi * kT + tx;
            if (k < n && i < n) Vt[(size_t)k * n + i] = t[tx][r];
        }
    }
    if (bi == 0 && threadIdx.y == 0 && col < n) sigma[col] = clamp ? fmax(d[src], 0.0) : d[src];
}

// out[c][r] = in[r][c] for one 32 x 32 tile.
__global__ __launch_bounds__(kT * kRows) void large_transpose_kernel(const double* in, int rows, int cols,
                                                                    double* out) {
    __shared__ double t[kT][kT + 1];
    const int br = blockIdx.y, bc = blockIdx.x, tx = threadIdx.x;
    for (int r = threadIdx.y; r < kT; r += kRows) {
        const int i = br * kT + r, j = bc * kT + tx;
        t[r][tx] = (i < rows && j < cols) ? in[(size_t)i * cols + j] : 0.0;
    }
    __syncthreads();
    for (int r = threadIdx.y; r < kT; r += kRows) {
        const int j = bc * kT + r, i = br * kT + tx;
        if (i < rows && j < cols) out[(size_t)j * rows + i] = t[tx][r];
    }
}

}  // namespace

hipError_t large_transpose(const double* in, int rows, int cols, double* out, hipStream_t st) {
    hipLaunchKernelGGL(large_transpose_kernel, dim3((cols + kT - 1) / kT, (rows + kT - 1) / kT), dim3(kT, kRows), 0,
                       st, in, rows, cols, out);
    return hipGetLastError();
}

hipError_t large_symmetrize(double* P, int n, hipStream_t st) {
    const int nt = (n + kT - 1) / kT;
    hipLaunchKernelGGL(large_sym_kernel, dim3(nt, nt), dim3(kT, kRows), 0, st, P, n);
    return hipGetLastError();
}

hipError_t large_rank(const double* d, int n, int* perm, int* bad, hipStream_t st) {
    if (n > kLargeMaxN) return hipErrorInvalidValue;
    hipLaunchKernelGGL(large_rank_kernel, dim3(1), dim3(1024), 0, st, d, n, perm, bad);
    return hipGetLastError();
}

hipError_t large_permute(const double* V, const double* d, const int* perm, int n, double* Vs, double* Vt,
                         double* sigma, bool clamp, hipStream_t st) {
    const int nt = (n + kT - 1) / kT;
    hipLaunchKernelGGL(large_permute_kernel, dim3(nt, nt), dim3(kT, kRows), 0, st, V, d, perm, n, Vs, Vt, sigma,
                       clamp ? 1 : 0);
    return hipGetLastError();
}

}  // namespace lsvd
