// lsvd_polish.hpp -- launch interface of the batched square GEMM used by
// the digital-polish entry point (polish.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lsvd {

// One n x n product of the batch: C = op(A) op(B), offsets in doubles.
struct BgemmItem {
    int64_t a_off, b_off, c_off;
    int n;
    int pad;
};

// ta / tb: use the transpose of A / B.  max_n sizes the grid.
hipError_t launch_bgemm(const BgemmItem* items, int count, int max_n, const double* a, const double* b, double* c,
                        bool ta, bool tb, hipStream_t st);

}  // namespace lsvd
