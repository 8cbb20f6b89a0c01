// givens.hip -- rotor-stream export: batched Givens elimination of
// orthogonal matrices into rotation events, the computation of
// HardwareSchedule::from_orthogonal_matrix (src/lie_svd_compiler.rs:144-197).
//
// One 64-lane wavefront per matrix.  Lane l owns columns l, l + 64, ... of
// the working copy R (row-major d x d, in LDS for d <= 128, otherwise in a
// global slot).  The eliminations run in the reference's order (column j,
// then rows i = j+1..d-1) with its arithmetic -- norm = sqrt(a a + b b),
// c = a / norm, s = b / norm, r_j <- c r_j + s r_i, r_i <- -s r_j + c r_i,
// no contraction into FMAs -- so c, s and the signs are the reference's
// bit for bit.  The pivot row j stays in registers for its whole pass (row
// j is final after it); row i is read once and written once per pass,
// prefetched one step ahead together with the broadcast r_ij.  Every lane
// carries the pivot a = r_jj itself (a <- c a + s b, the same operation
// the owner lane applies to its copy), so the only sequential chain per
// rotation is a, b -> norm -> c, s -> a.  Events are staged one per lane
// and stored 64 at a time; theta = atan2(s, c) is computed at that flush.
#include <hip/hip_runtime.h>

#include "lsvd_schedule.hpp"

namespace lsvd {
namespace {

constexpr int kLanes = 64;

template <int CPL>
__device__ __forceinline__ void flush_events(const GivensLaunch& L, int64_t base, int lane, int valid, int sij,
                                             double sc, double ss) {
    if (lane < valid) {
        const int64_t e = base + lane;
        L.ev_ij[2 * e] = sij & 0xffff;
        L.ev_ij[2 * e + 1] = sij >> 16;
        L.ev_cs[2 * e] = sc;
        L.ev_cs[2 * e + 1] = ss;
        if (L.ev_theta) L.ev_theta[e] = atan2(ss, sc);
    }
}

template <int CPL, bool IN_LDS>
__global__ __launch_bounds__(kLanes) void givens_kernel(GivensLaunch L) {
#pragma clang fp contract(off)
    extern __shared__ double lds[];
    const int lane = threadIdx.x;
    double* R = IN_LDS ? lds : L.scratch + (int64_t)blockIdx.x * L.scratch_stride;
    for (int b = blockIdx.x; b < L.count; b += gridDim.x) {
        const GivensItem it = L.items[b];
        const int d = it.d;
        const double* src = L.pool + it.src;
        __syncthreads();  // the previous matrix's last reads of R
        // working copy: coalesced source rows; a transposed source lands in
        // other lanes' columns, hence the barrier below
        for (int r = 0; r < d; ++r) {
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                const int col = lane + kLanes * c;
                if (col < d) {
                    const double v = src[(int64_t)r * d + col];
                    if (L.transpose) R[(int64_t)col * d + r] = v;
                    else R[(int64_t)r * d + col] = v;
                }
            }
        }
        int ne = 0;
        int sij = 0;
        double sc = 0.0, ss = 0.0;
        for (int j = 0; j < d; ++j) {
            // column j of every row was last written by its owner lane in
            // the previous pass; the broadcast loads below read it
            __syncthreads();
            const int64_t rowj = (int64_t)j * d;
            double rj[CPL], nx[CPL];
#pragma unroll
            for (int c = 0; c < CPL; ++c) {
                const int col = lane + kLanes * c;
                rj[c] = col < d ? R[rowj + col] : 0.0;
                nx[c] = 0.0;
            }
            double a = R[rowj + j];
            double nb = 0.0;
            if (j + 1 < d) {
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    const int col = lane + kLanes * c;
                    nx[c] = col < d ? R[rowj + d + col] : 0.0;
                }
                nb = R[rowj + d + j];
            }
            for (int i = j + 1; i < d; ++i) {
                double ri[CPL];
#pragma unroll
                for (int c = 0; c < CPL; ++c) ri[c] = nx[c];
                const double bv = nb;
                const int64_t rowi = (int64_t)i * d;
                if (i + 1 < d) {
#pragma unroll
                    for (int c = 0; c < CPL; ++c) {
                        const int col = lane + kLanes * c;
                        nx[c] = col < d ? R[rowi + d + col] : 0.0;
                    }
                    nb = R[rowi + d + j];
                }
                if (fabs(bv) <= 1e-300) continue;  // (:158-160), uniform
                const double norm = sqrt(a * a + bv * bv);
                const double cs = a / norm;
                const double sn = bv / norm;
#pragma unroll
                for (int c = 0; c < CPL; ++c) {
                    if (kLanes * c + kLanes - 1 < j) continue;  // columns < j no longer matter
                    const int col = lane + kLanes * c;
                    const double t = cs * rj[c] + sn * ri[c];
                    const double u = -sn * rj[c] + cs * ri[c];
                    rj[c] = t;
                    if (col < d) R[rowi + col] = u;
                }
                a = cs * a + sn * bv;
                if (lane == (ne & (kLanes - 1))) {
                    sij = i | (j << 16);
                    sc = cs;
                    ss = sn;
                }
                ++ne;
                if ((ne & (kLanes - 1)) == 0) flush_events<CPL>(L, it.ev + ne - kLanes, lane, kLanes, sij, sc, ss);
            }
            if (lane == 0) L.signs[it.sg + j] = isnan(a) ? a : copysign(1.0, a);
        }
        const int rem = ne & (kLanes - 1);
        if (rem) flush_events<CPL>(L, it.ev + (ne - rem), lane, rem, sij, sc, ss);
        if (lane == 0) L.n_events[it.idx] = ne;
    }
}

template <int CPL, bool IN_LDS>
hipError_t launch_one(const GivensLaunch& L, size_t lds, int grid, hipStream_t st) {
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&givens_kernel<CPL, IN_LDS>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((givens_kernel<CPL, IN_LDS>), dim3(grid), dim3(kLanes), lds, st, L);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_givens(const GivensLaunch& L, int d_max, int grid, hipStream_t st) {
    if (L.count <= 0) return hipSuccess;
    if (d_max <= kGivensLdsMaxD) {
        const size_t lds = sizeof(double) * (size_t)d_max * (size_t)d_max;
        if (d_max <= 64) return launch_one<1, true>(L, lds, grid, st);
        return launch_one<2, true>(L, lds, grid, st);
    }
    if (d_max <= 256) return launch_one<4, false>(L, 0, grid, st);
    if (d_max <= 512) return launch_one<8, false>(L, 0, grid, st);
    if (d_max <= kGivensMaxD) return launch_one<16, false>(L, 0, grid, st);
    return hipErrorInvalidValue;
}

}  // namespace lsvd
