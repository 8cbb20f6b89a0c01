// lsvd_device.hpp -- device-side building blocks shared by the gfx950 kernels.
//
// Everything here runs inside one workgroup: block reductions over 64-wide
// wavefronts, a register-blocked in-workgroup GEMM over generic (flat)
// pointers that may point at LDS or at global memory, and the round-robin
// pair schedule.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lsvd_types.hpp"

namespace lsvd {

constexpr int kWave = 64;

// Thread index through an opaque move.  The batched kernels loop over work
// items inside one launch; reading the index this way makes every address
// derived from it a per-item value, so the compiler recomputes it where it
// is used instead of hoisting all phases' addresses out of the item loop
// (which keeps them live across the whole kernel and forces spills).
__device__ __forceinline__ int tix() {
    int t;
    asm volatile("v_mov_b32 %0, %1" : "=v"(t) : "v"((int)threadIdx.x));
    return t;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
    return v;
}

__device__ __forceinline__ int wave_or(int v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v |= __shfl_xor(v, off, kWave);
    return v;
}

// Block-wide sum.  `red` is LDS scratch of at least NT/64 doubles.  Every
// thread gets the result.  Contains two barriers.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red) {
    v = wave_sum(v);
    const int w = tix() / kWave;
    if ((tix() & (kWave - 1)) == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < NT / kWave; ++i) s += red[i];
    __syncthreads();
    return s;
}

// Two block-wide sums with the barriers of one; `red` holds >= 2 NT/64.
template <int NT>
__device__ __forceinline__ void block_sum2(double& a, double& b, double* red) {
    a = wave_sum(a);
    b = wave_sum(b);
    const int w = tix() / kWave;
    if ((tix() & (kWave - 1)) == 0) {
        red[2 * w] = a;
        red[2 * w + 1] = b;
    }
    __syncthreads();
    double sa = 0.0, sb = 0.0;
#pragma unroll
    for (int i = 0; i < NT / kWave; ++i) {
        sa += red[2 * i];
        sb += red[2 * i + 1];
    }
    __syncthreads();
    a = sa;
    b = sb;
}

template <int NT>
__device__ __forceinline__ double block_max(double v, double* red) {
    v = wave_max(v);
    const int w = tix() / kWave;
    if ((tix() & (kWave - 1)) == 0) red[w] = v;
    __syncthreads();
    double s = red[0];
#pragma unroll
    for (int i = 1; i < NT / kWave; ++i) s = fmax(s, red[i]);
    __syncthreads();
    return s;
}

template <int NT>
__device__ __forceinline__ int block_or(int v, double* red) {
    v = wave_or(v);
    int* ri = reinterpret_cast<int*>(red);
    const int w = tix() / kWave;
    if ((tix() & (kWave - 1)) == 0) ri[w] = v;
    __syncthreads();
    int s = 0;
#pragma unroll
    for (int i = 0; i < NT / kWave; ++i) s |= ri[i];
    __syncthreads();
    return s;
}

// Row-major view with optional transpose: (i, j) -> p[i*ld + j], or
// p[j*ld + i] when transposed.
struct View {
    const double* p;
    int ld;
    int trans;
    __device__ __forceinline__ double operator()(int i, int j) const {
        return trans ? p[(size_t)j * ld + i] : p[(size_t)i * ld + j];
    }
};

// Upper-triangular view of the R factor stored in LAPACK compact form
// (Householder vectors below the diagonal are read as zero).
struct UpperView {
    const double* p;
    int ld;
    __device__ __forceinline__ double operator()(int i, int j) const {
        return i <= j ? p[(size_t)i * ld + j] : 0.0;
    }
};

// In-workgroup GEMM: for every (i, j) in M x N, epi(i, j, sum_k a(i,k) b(k,j)).
// Each thread owns TM x TN output micro-tiles; consecutive threads take
// consecutive column tiles so B-row reads coalesce and A reads broadcast.
template <int NT, int TM, int TN, class LA, class LB, class Epi>
__device__ __forceinline__ void wg_gemm_t(int M, int N, int K, const LA& la, const LB& lb, Epi&& epi) {
    const int tm = (M + TM - 1) / TM;
    const int tn = (N + TN - 1) / TN;
    for (int t = tix(); t < tm * tn; t += NT) {
        const int i0 = (t / tn) * TM;
        const int j0 = (t % tn) * TN;
        double acc[TM][TN];
#pragma unroll
        for (int a = 0; a < TM; ++a)
#pragma unroll
            for (int b = 0; b < TN; ++b) acc[a][b] = 0.0;
        for (int k = 0; k < K; ++k) {
            double av[TM], bv[TN];
#pragma unroll
            for (int a = 0; a < TM; ++a) av[a] = (i0 + a < M) ? la(i0 + a, k) : 0.0;
#pragma unroll
            for (int b = 0; b < TN; ++b) bv[b] = (j0 + b < N) ? lb(k, j0 + b) : 0.0;
#pragma unroll
            for (int a = 0; a < TM; ++a)
#pragma unroll
                for (int b = 0; b < TN; ++b) acc[a][b] = fma(av[a], bv[b], acc[a][b]);
        }
#pragma unroll
        for (int a = 0; a < TM; ++a)
#pragma unroll
            for (int b = 0; b < TN; ++b)
                if (i0 + a < M && j0 + b < N) epi(i0 + a, j0 + b, acc[a][b]);
    }
}

// Picks the micro-tile from the output size so that there is roughly one
// tile per thread or more.
template <int NT, class LA, class LB, class Epi>
__device__ __forceinline__ void wg_gemm(int M, int N, int K, const LA& la, const LB& lb, Epi&& epi) {
    const int cells = M * N;
    if (cells >= 16 * NT)
        wg_gemm_t<NT, 4, 4>(M, N, K, la, lb, epi);
    else if (cells >= 4 * NT)
        wg_gemm_t<NT, 2, 2>(M, N, K, la, lb, epi);
    else
        wg_gemm_t<NT, 1, 1>(M, N, K, la, lb, epi);
}

// Circle-method round-robin schedule (the reference's layer_pairs,
// src/lie_svd_joint.rs:524-539): slot k of `layer` for n indices.  Returns
// (p, q) with p < q; q == -1 marks a slot whose partner is the phantom
// index of odd n (p is then unpaired in this layer).
__device__ __forceinline__ void rr_pair(int n, int layer, int k, int& p, int& q) {
    const int m = n + (n & 1);
    const int ring = m - 1;
    int a, b;
    if (k == 0) {
        a = m - 1;
        b = layer % ring;
    } else {
        a = (layer + k) % ring;
        b = (layer + ring - k) % ring;
    }
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    if (hi >= n) {
        p = lo;
        q = -1;
    } else {
        p = lo;
        q = hi;
    }
}

__device__ __forceinline__ int rr_layers(int n) { return (n & 1) ? n : n - 1; }

// Jacobi rotation of one pair (src/lie_svd_small.rs:217-243) from its 2x2
// diagonal block in slot order (x00 at the slot's first index; fwd: that
// index is the reference's p < q); c, s and t are oriented to the slot
// order.  The reference's theta = d / (2 a_pq), t = sgn(theta) /
// (|theta| + sqrt(theta^2 + 1)), c = 1 / sqrt(t^2 + 1), s = t c are
// evaluated as t = sgn |2 a_pq| / (|d| + r), c = (|d| + r) / h,
// s = sgn |2 a_pq| / h with r = sqrt(d^2 + 4 a_pq^2), h = sqrt((|d| + r)^2
// + 4 a_pq^2): the same quantities (within a few ulps) with a shorter
// dependent chain (one sqrt, then a division and a reciprocal square root
// side by side).  d and 2 a_pq are first scaled by the same power of two, so
// the squares cannot overflow or underflow.
__device__ __forceinline__ int jacobi_angle(double x00, double x11, double apq, bool fwd, double& c, double& sv,
                                            double& tv) {
    const bool rot = fabs(apq) >= 1e-300;
    const double app = fwd ? x00 : x11, aqq = fwd ? x11 : x00;
    const double d = aqq - app, a2 = 2.0 * apq;
    const int ex = __builtin_amdgcn_frexp_exp(fmax(fabs(d), fabs(a2)));
    const double ds = ldexp(d, -ex), as = fabs(ldexp(a2, -ex));
    const double r = sqrt(fma(ds, ds, as * as));
    const double den = fabs(ds) + r;
    const double sgn = (d == 0.0 || ((d > 0.0) == (a2 > 0.0))) ? 1.0 : -1.0;  // sgn(theta), theta = +-0 -> +
    const double tt = sgn * as / den;
    const double ih = rsqrt(fma(den, den, as * as));
    c = rot ? den * ih : 1.0;
    const double sn = rot ? sgn * as * ih : 0.0;
    sv = fwd ? sn : -sn;
    tv = rot ? (fwd ? tt : -tt) : 0.0;
    return rot ? 1 : 0;
}

}  // namespace lsvd
