// lsvd_tensor.hpp -- launch interface of the HO-SVD tensor kernels
// (tensor.hip) used by lsvd_tensor.cpp.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lsvd {

constexpr int kTensorMaxN = 64;  // largest extent of one mode

// Mode Grams of `count` tensors (n1 x n2 x n3 each, contiguous): G of tensor
// t, mode m at g + g_off[3 t + m] (dims[m]^2 doubles).
hipError_t tensor_mode_grams(const double* x, int count, int n1, int n2, int n3, double* g, const int64_t* g_off,
                             hipStream_t st);
// out = x x_mode op(Q) for every tensor: Q of tensor t at q + q_off[t *
// q_stride_mode + mode], r x dims[mode] (or dims[mode] x r with trans).
hipError_t tensor_mode_product(const double* x, int count, int n1, int n2, int n3, int mode, int r, const double* q,
                               const int64_t* q_off, int q_stride_mode, bool trans, double* out, hipStream_t st);
// off[t] = norm of the entries off the superdiagonal, ratio[t] = superdiagonal
// mass / total mass (either may be null).
hipError_t tensor_norms(const double* x, int count, int n1, int n2, int n3, double* off, double* ratio,
                        hipStream_t st);

}  // namespace lsvd
