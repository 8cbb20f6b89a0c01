// lsvd_polar.hpp -- launch interface of the large-matrix polar kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include "../../include/lie_svd_amd.h"

namespace lsvd {

// newton_schulz_polar on device buffers (dA, dQ: m x n, m >= n >= 1) on the
// handle's stream; synchronous.  Precision as lsvd_polar_f64.
int polar_run(lsvd_handle_t h, const double* dA, int m, int n, int precision, double* dQ, lsvd_polar_info_t* info);

enum PolarPrec { kPolF64 = 0, kPolF32 = 1, kPolBF16 = 2 };
enum Epilogue { kEpiStore = 0, kEpi3IMinus = 1, kEpiHalfDelta = 2, kEpiOrth = 3 };

// M = alpha I - X^T X (n x n) from X (m x n); X/M are f64 (kPolF64) or
// f32.  *orth_sq += ||X^T X - I||_F^2 when orth_sq is non-null.
hipError_t ns_gram(int prec, int m, int n, const void* X, void* M, double* orth_sq, hipStream_t st,
                   double alpha = 3.0);
// Xn = c X M; *dsq += ||Xn - X||_F^2; *bad |= any non-finite entry of Xn.
hipError_t ns_update(int prec, int m, int n, const void* X, const void* M, void* Xn, double* dsq, int* bad,
                     hipStream_t st, double c = 0.5);
// *sum += ||X^T X - I||_F^2 in FP64.
hipError_t ns_orth(int m, int n, const double* X, double* sum, hipStream_t st);
// C = alpha op(A) B, FP64; op(A) = A^T when trans_a (A stored K x M).
hipError_t gemm_f64(bool trans_a, int M, int N, int K, const double* A, const double* B, double* C, double alpha,
                    hipStream_t st);
// C = op(A) B with FP32 (kPolF32) or BF16-rounded (kPolBF16) operands, FP32 out.
hipError_t gemm_lowp(int prec, bool trans_a, int M, int N, int K, const float* A, const float* B, float* C,
                     hipStream_t st);
// *out = ||x||^2 (overwritten), deterministic; partials: kSumsqBlocks doubles.
constexpr int kSumsqBlocks = 256;
hipError_t polar_sumsq(const double* x, size_t n, double* out, double* partials, hipStream_t st);
// *out = ||B - B^T||_F^2 of an n x n matrix (overwritten), deterministic.
hipError_t polar_skew_sumsq(const double* b, int n, double* out, double* partials, hipStream_t st);
hipError_t polar_scale(const double* x, size_t n, const double* sumsq, void* y, bool f32, hipStream_t st);
hipError_t polar_widen(const float* x, size_t n, double* y, hipStream_t st);
hipError_t polar_zero(double* p, int n, hipStream_t st);

}  // namespace lsvd
