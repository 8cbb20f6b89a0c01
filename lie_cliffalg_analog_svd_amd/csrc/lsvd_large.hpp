// lsvd_large.hpp -- the chip-wide route of LieSvdSmall::solve for one large
// square matrix (large.hip kernels, lsvd_large.cpp driver).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/lie_svd_amd.h"

namespace lsvd {

constexpr int kLargeMaxN = 1024;          // the single-matrix API's limit
constexpr int kLargeMinNDefault = 129;    // smallest n routed chip-wide by default (above the on-chip sizes)

// Block Jacobi eigen-decomposition of one symmetric n x n matrix
// (block_jacobi.hip).  A and V are ld x ld (ld = n padded to whole tiles);
// a layer's (nb / 2) bb x bb block rotations go to a W buffer of
// (nb / 2) bb^2 doubles.
struct BlockJacobi {
    int n, ld, nb, bb;
    double* A;
    double* V;
    double* partials;  // kBjOffBlocks doubles
    int* rotations;
};
constexpr int kBjOffBlocks = 256;
int bj_tile_size(int n);
hipError_t bj_setup(const BlockJacobi& j, const double* P, hipStream_t st);
// One layer: inner (A -> W), then A <- G^T A G and V <- V G in one launch
// (reads W).
hipError_t bj_inner(const BlockJacobi& j, int layer, double* W, hipStream_t st);
hipError_t bj_update(const BlockJacobi& j, int layer, const double* W, hipStream_t st);
// *out = sum of squares of the upper off-diagonal entries (overwritten).
hipError_t bj_offnorm(const BlockJacobi& j, double* out, hipStream_t st);
hipError_t bj_extract(const BlockJacobi& j, double* d, double* Vc, hipStream_t st);

hipError_t large_symmetrize(double* P, int n, hipStream_t st);
hipError_t large_rank(const double* d, int n, int* perm, int* bad, hipStream_t st);
// Vt may be null; clamp: sigma = max(d, 0) (the SVD) instead of d (eigh).
hipError_t large_permute(const double* V, const double* d, const int* perm, int n, double* Vs, double* Vt,
                         double* sigma, bool clamp, hipStream_t st);

// out (cols x rows) = in^T (in: rows x cols), through LDS tiles.
hipError_t large_transpose(const double* in, int rows, int cols, double* out, hipStream_t st);

// LieSvdSmall::solve (m == n) or solve_rectangular of one m x n matrix with
// every step spread over the chip: u m x k, s k, vt k x n (k = min(m, n));
// a tall matrix runs the polar route on itself (U = Q V is m x k), a wide
// one is solved transposed.  Pointers follow the handle's pointer mode.
// Sets *done = false (and returns LSVD_OK) when the polar factor is not
// trusted or U = Q V fails the orthogonality test: those cases take the
// reference's repair branch, which the caller runs through the
// workgroup-per-matrix kernels.
int svd_large(lsvd_handle_t h, const double* a, int m, int n, double* u, double* s, double* vt, lsvd_info_t* info,
              bool* done);

// eigh_jacobi_full (src/lie_svd_small.rs:212-285) of one symmetric n x n
// matrix by the block Jacobi method: v (n x n, eigenvectors as columns)
// and eig (descending).  Pointers follow the handle's pointer mode; sets
// *done = false when the diagonal holds a non-finite value (the caller then
// runs the workgroup kernels).
int eigh_large(lsvd_handle_t h, const double* p, int n, double* v, double* eig, lsvd_info_t* info, bool* done);

}  // namespace lsvd
