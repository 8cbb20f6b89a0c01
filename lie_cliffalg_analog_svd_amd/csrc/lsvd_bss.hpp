// lsvd_bss.hpp -- launch interface of the Phase-BSS kernels (bss.hip) used by
// lsvd_bss.cpp.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lsvd {

constexpr int kBssMaxChannels = 64;  // channels of one observation block
constexpr int kBssFillBlocks = 8192;  // cross-product workgroups aimed for

// Sample splits of the cross-products of `count` blocks at `nq` lags.
int bss_splits(int count, int ch, int S, int nq);

// mean[p ch + i] = mean over t of x[p][i][t] (count blocks of ch x S).
hipError_t bss_means(const double* x, int count, int ch, int S, double* mean, hipStream_t st);
// Lagged cross-products of every block: for lag lags[q] (q < nq) and
// a = b = x - mean (mean may be null), the sum over t < S - lag of
// a[i][t + lag] b[j][t], split over `nsplit` spans of t into
// part[((p nq + q) nsplit + split) ch^2 + i ch + j].
hipError_t bss_cross(const double* x, const double* mean, int count, int ch, int S, const int* lags, int nq,
                     int nsplit, double* part, hipStream_t st);
// out[(p nq + q) ch^2 + i ch + j] = c_ij (sym: (c_ij + c_ji) / 2), where c is
// the split sum of part (in split order) divided by denom[q].
hipError_t bss_reduce(const double* part, int count, int ch, int nq, int nsplit, const double* denom, bool sym,
                      double* out, hipStream_t st);
// W[p] (ch x ch) from u[p], sigma[p]: floor = eps max(max sigma, 1),
// W[i][j] = u[j][i] / sqrt(max(sigma_i, floor)); floor[p] may be null.
hipError_t bss_whitening(const double* u, const double* sigma, int count, int ch, double eps, double* W,
                         double* floor, hipStream_t st);
// out[p] = op(M[p]) (in[p] - mean[p]) with op(M) = M or M^T; mean may be null.
hipError_t bss_apply(const double* M, bool trans, const double* in, const double* mean, int count, int ch, int S,
                     double* out, hipStream_t st);
// out[p] = V[p]^T W[p] (ch x ch).
hipError_t bss_unmixing(const double* V, const double* W, int count, int ch, double* out, hipStream_t st);
// coh[p ch + i] = channel phase coherence of row i of y[p].
hipError_t bss_coherence(const double* y, int count, int ch, int S, double* coh, hipStream_t st);

}  // namespace lsvd
