// tensor.hip -- the tensor steps of LieSvdTensor::hosvd3
// (src/lie_svd_tensor.rs:31-62) for a batch of 3-way tensors of one shape
// n1 x n2 x n3 (row-major, the last index fastest):
//   mode Gram  G_m = X_(m) X_(m)^T           (mode_gram, :73-119)
//   mode product  Y = X x_m Q (or Q^T)       (rotate_tensor3, :121-145, as
//                 three successive mode products instead of one 6-deep sum)
//   trace norms   off-superdiagonal norm, superdiagonal mass (:147-183)
// All are small dense contractions (n <= 64 per mode), run as plain FP64
// loops: one workgroup per (tensor, mode) Gram with the operand chunk staged
// in LDS, one thread per output entry for the mode products.
#include <hip/hip_runtime.h>
#include <math.h>

#include "lsvd_device.hpp"
#include "lsvd_tensor.hpp"

namespace lsvd {
namespace {

constexpr int kGramChunk = 64;  // "other" index columns staged per step

// x[i0][i1][i2] with the mode index `a` in slot `mode` and the flattened
// other two indices `o` (the lower slot's index major).
__device__ __forceinline__ size_t t_index(int mode, int a, int o, int n1, int n2, int n3) {
    if (mode == 0) return (size_t)a * n2 * n3 + o;                 // o = j n3 + k
    if (mode == 1) return ((size_t)(o / n3) * n2 + a) * n3 + o % n3;  // o = i n3 + k
    return (size_t)o * n3 + a;                                        // o = i n2 + j
}

// G (n x n, n = dims[mode]) of tensor blockIdx.x, mode blockIdx.y.  Each
// entry is summed over the other two indices in the reference's loop order,
// one rounding per product and per add (no FMA contraction), so G has the
// reference's bits.
__global__ __launch_bounds__(256) void mode_gram_kernel(const double* x, int n1, int n2, int n3, double* g,
                                                        const int64_t* g_off) {
#pragma clang fp contract(off)
    __shared__ double sx[kTensorMaxN * (kGramChunk + 1)];
    const int t = blockIdx.x, mode = blockIdx.y;
    const int dims[3] = {n1, n2, n3};
    const int n = dims[mode];
    const int other = n1 * n2 * n3 / n;
    const double* xt = x + (size_t)t * n1 * n2 * n3;
    double* gt = g + g_off[(size_t)t * 3 + mode];
    constexpr int L = kGramChunk + 1;
    const int pairs = n * n;
    // each thread keeps up to 16 accumulators (n <= 64: 4096 entries / 256)
    double acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0.0;
    for (int o0 = 0; o0 < other; o0 += kGramChunk) {
        const int cw = other - o0 < kGramChunk ? other - o0 : kGramChunk;
        for (int e = threadIdx.x; e < n * kGramChunk; e += blockDim.x) {
            const int a = e / kGramChunk, c = e % kGramChunk;
            sx[a * L + c] = c < cw ? xt[t_index(mode, a, o0 + c, n1, n2, n3)] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int e = threadIdx.x + q * 256;
            if (e >= pairs) break;
            const int a = e / n, b = e % n;
            double s = acc[q];
            for (int c = 0; c < cw; ++c) s += sx[a * L + c] * sx[b * L + c];
            acc[q] = s;
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int e = threadIdx.x + q * 256;
        if (e >= pairs) break;
        gt[e] = acc[q];
    }
}

// out = x x_mode op(Q): out[.. a ..] = sum_i op(Q)[a][i] x[.. i ..], where
// op(Q) = Q (r x n) or Q^T (Q stored n x r, trans = 1).  The output has the
// mode's extent r; q_off gives each tensor's Q.
__global__ void mode_product_kernel(const double* x, int n1, int n2, int n3, int mode, int r, const double* q,
                                    const int64_t* q_off, int q_stride_mode, int trans, double* out, int count) {
    const int dims[3] = {n1, n2, n3};
    int od[3] = {n1, n2, n3};
    od[mode] = r;
    const size_t in_sz = (size_t)n1 * n2 * n3, out_sz = (size_t)od[0] * od[1] * od[2];
    const int n = dims[mode];
    const size_t total = out_sz * count;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int t = (int)(e / out_sz);
        const size_t w = e % out_sz;
        const int c0 = (int)(w / ((size_t)od[1] * od[2])), c1 = (int)((w / od[2]) % od[1]), c2 = (int)(w % od[2]);
        const int cc[3] = {c0, c1, c2};
        const int a = cc[mode];
        const double* xt = x + (size_t)t * in_sz;
        const double* qt = q + q_off[(size_t)t * q_stride_mode + mode];
        double s = 0.0;
        for (int i = 0; i < n; ++i) {
            int ii[3] = {c0, c1, c2};
            ii[mode] = i;
            const double qv = trans ? qt[(size_t)i * r + a] : qt[(size_t)a * n + i];
            s += qv * xt[((size_t)ii[0] * n2 + ii[1]) * n3 + ii[2]];
        }
        out[(size_t)t * out_sz + w] = s;
    }
}

// Per tensor: off-superdiagonal norm and superdiagonal mass ratio (:147-183).
__global__ __launch_bounds__(256) void tensor_norms_kernel(const double* x, int n1, int n2, int n3, double* off,
                                                           double* ratio) {
    __shared__ double red[8];
    const int t = blockIdx.x;
    const size_t sz = (size_t)n1 * n2 * n3;
    const double* xt = x + t * sz;
    double od = 0.0, diag = 0.0;
    for (size_t e = threadIdx.x; e < sz; e += blockDim.x) {
        const int i = (int)(e / ((size_t)n2 * n3)), j = (int)((e / n3) % n2), k = (int)(e % n3);
        const double v = xt[e];
        if (i == j && j == k) diag += v * v;
        else od += v * v;
    }
    block_sum2<256>(od, diag, red);
    if (threadIdx.x == 0) {
        if (off) off[t] = sqrt(od);
        if (ratio) ratio[t] = diag / fmax(od + diag, 1e-300);
    }
}

}  // namespace

hipError_t tensor_mode_grams(const double* x, int count, int n1, int n2, int n3, double* g, const int64_t* g_off,
                             hipStream_t st) {
    hipLaunchKernelGGL(mode_gram_kernel, dim3(count, 3), dim3(256), 0, st, x, n1, n2, n3, g, g_off);
    return hipGetLastError();
}

hipError_t tensor_mode_product(const double* x, int count, int n1, int n2, int n3, int mode, int r, const double* q,
                               const int64_t* q_off, int q_stride_mode, bool trans, double* out, hipStream_t st) {
    const int dims[3] = {n1, n2, n3};
    const size_t out_sz = (size_t)n1 * n2 * n3 / dims[mode] * r * count;
    const int blocks = (int)((out_sz + 255) / 256 < 8192 ? (out_sz + 255) / 256 : 8192);
    hipLaunchKernelGGL(mode_product_kernel, dim3(blocks > 0 ? blocks : 1), dim3(256), 0, st, x, n1, n2, n3, mode, r,
                       q, q_off, q_stride_mode, trans ? 1 : 0, out, count);
    return hipGetLastError();
}

hipError_t tensor_norms(const double* x, int count, int n1, int n2, int n3, double* off, double* ratio,
                        hipStream_t st) {
    hipLaunchKernelGGL(tensor_norms_kernel, dim3(count), dim3(256), 0, st, x, n1, n2, n3, off, ratio);
    return hipGetLastError();
}

}  // namespace lsvd
