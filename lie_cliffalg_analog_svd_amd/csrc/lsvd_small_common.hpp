// lsvd_small_common.hpp -- per-matrix building blocks shared by the batched
// LieSvdSmall kernels: Gram-Schmidt with the reference's rank-deficient
// fallback, Householder QR in compact form, and the application of its Q.
// All run inside one workgroup of NT threads on flat pointers (LDS or
// global); `red` is LDS reduction scratch, `scratch` holds >= N doubles.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include "lsvd_device.hpp"

namespace lsvd {

// gram_schmidt_orthonormal (src/lie_svd_small.rs:82-144): MGS over the N
// columns (length M >= N) of `a` into `out` (M x N, leading dimension ld).
// Rank-deficient columns take the standard-basis direction e_k (k < M) with
// the largest residual; its norm is evaluated as 1 - sum_c out[k][c]^2,
// which equals the reference's MGS projection of e_k up to rounding.
template <int NT, class V>
__device__ void gram_schmidt_g(int M, int N, const V& a, double* out, int ld, double* red) {
    for (int j = 0; j < N; ++j) {
        for (int i = tix(); i < M; i += NT) out[(size_t)i * ld + j] = a(i, j);
        __syncthreads();
        for (int k = 0; k < j; ++k) {
            double d = 0.0;
            for (int i = tix(); i < M; i += NT) d += out[(size_t)i * ld + k] * out[(size_t)i * ld + j];
            d = block_sum<NT>(d, red);
            for (int i = tix(); i < M; i += NT) out[(size_t)i * ld + j] -= d * out[(size_t)i * ld + k];
            __syncthreads();
        }
        double ns = 0.0;
        for (int i = tix(); i < M; i += NT) ns += out[(size_t)i * ld + j] * out[(size_t)i * ld + j];
        const double norm = sqrt(block_sum<NT>(ns, red));
        if (norm >= 1e-10) {
            for (int i = tix(); i < M; i += NT) out[(size_t)i * ld + j] /= norm;
            __syncthreads();
            continue;
        }
        // Best standard-basis candidate: thread k evaluates e_k.
        double best = -1.0;
        int bestk = 0;
        for (int k = tix(); k < M; k += NT) {
            double r = 1.0;
            for (int c = 0; c < j; ++c) r -= out[(size_t)k * ld + c] * out[(size_t)k * ld + c];
            const double en = sqrt(fmax(r, 0.0));
            if (en > best) {
                best = en;
                bestk = k;
            }
        }
        // argmax with lowest index on ties (the reference keeps the first max)
        double* rd = red;
        __syncthreads();
        if (tix() == 0) {
            rd[0] = -1.0;
            reinterpret_cast<int*>(rd)[2] = 0;
        }
        __syncthreads();
        for (int w = 0; w < NT; ++w) {  // serialised, fallback path only
            if (tix() == w && best > rd[0]) {
                rd[0] = best;
                reinterpret_cast<int*>(rd)[2] = bestk;
            }
            __syncthreads();
        }
        const double bn = rd[0];
        const int bk = reinterpret_cast<int*>(rd)[2];
        __syncthreads();
        if (bn > 0.0 && isfinite(bn)) {
            for (int i = tix(); i < M; i += NT) out[(size_t)i * ld + j] = (i == bk) ? 1.0 : 0.0;
            __syncthreads();
            for (int k = 0; k < j; ++k) {
                double d = 0.0;
                for (int i = tix(); i < M; i += NT) d += out[(size_t)i * ld + k] * out[(size_t)i * ld + j];
                d = block_sum<NT>(d, red);
                for (int i = tix(); i < M; i += NT) out[(size_t)i * ld + j] -= d * out[(size_t)i * ld + k];
                __syncthreads();
            }
            double ns2 = 0.0;
            for (int i = tix(); i < M; i += NT) ns2 += out[(size_t)i * ld + j] * out[(size_t)i * ld + j];
            const double nn = sqrt(block_sum<NT>(ns2, red));
            for (int i = tix(); i < M; i += NT) out[(size_t)i * ld + j] /= nn;
        } else {
            for (int i = tix(); i < M; i += NT) out[(size_t)i * ld + j] = (i == j) ? 1.0 : 0.0;
        }
        __syncthreads();
    }
}

// Householder QR of the M x N matrix in `a` (ld), in place (LAPACK compact
// form: R on and above the diagonal, reflector tails below, tau[j]).
template <int NT>
__device__ void householder_qr_g(int M, int N, double* a, int ld, double* red, double* tau_s, double* scratch) {
    for (int j = 0; j < N; ++j) {
        double xs = 0.0;
        for (int i = j + 1 + tix(); i < M; i += NT) xs += a[(size_t)i * ld + j] * a[(size_t)i * ld + j];
        const double xnorm2 = block_sum<NT>(xs, red);
        const double alpha = a[(size_t)j * ld + j];
        double tau = 0.0, beta = alpha, scal = 0.0;
        if (xnorm2 > 0.0) {
            beta = -copysign(sqrt(alpha * alpha + xnorm2), alpha);
            tau = (beta - alpha) / beta;
            scal = 1.0 / (alpha - beta);
        }
        if (tix() == 0) tau_s[j] = tau;
        for (int i = j + 1 + tix(); i < M; i += NT) a[(size_t)i * ld + j] *= scal;
        __syncthreads();
        if (tau != 0.0) {
            // w_c = v^T a[:, c] for c > j, then a[:, c] -= tau v w_c
            for (int c = j + 1 + tix(); c < N; c += NT) {
                double w = a[(size_t)j * ld + c];
                for (int i = j + 1; i < M; ++i) w += a[(size_t)i * ld + j] * a[(size_t)i * ld + c];
                a[(size_t)j * ld + c] -= tau * w;
                scratch[c] = tau * w;  // scratch: tau * w_c for the trailing update
            }
            __syncthreads();
            const int cols = N - j - 1;
            for (int e = tix(); e < (M - j - 1) * cols; e += NT) {
                const int i = j + 1 + e / cols, c = j + 1 + e % cols;
                a[(size_t)i * ld + c] -= a[(size_t)i * ld + j] * scratch[c];
            }
        }
        if (tix() == 0) a[(size_t)j * ld + j] = beta;
        __syncthreads();
    }
}

// Y (M x N at y, ld) <- Q Y, Q = H_0 H_1 ... H_{N-1} from the compact form
// in global memory (qs, row-major M x N, ld N) and tau.
template <int NT>
__device__ void apply_q_g(int M, int N, const double* qs, double* y, int ld, const double* tau_s, double* scratch) {
    for (int j = N - 1; j >= 0; --j) {
        const double tau = tau_s[j];
        if (tau == 0.0) continue;  // uniform: tau is in LDS
        // w_c = y[j][c] + sum_{i>j} v_i y[i][c];  y[i][c] -= tau v_i w_c
        for (int c = tix(); c < N; c += NT) {
            double w = y[(size_t)j * ld + c];
            for (int i = j + 1; i < M; ++i) w += qs[(size_t)i * N + j] * y[(size_t)i * ld + c];
            scratch[c] = tau * w;
        }
        __syncthreads();
        for (int e = tix(); e < (M - j) * N; e += NT) {
            const int i = j + e / N, c = e % N;
            const double v = (i == j) ? 1.0 : qs[(size_t)i * N + j];
            y[(size_t)i * ld + c] -= v * scratch[c];
        }
        __syncthreads();
    }
}


}  // namespace lsvd
