// micro_svd.hip -- LieSvdMicro for n <= 4 on gfx950: one lane per matrix,
// everything in registers (src/lie_svd_micro.rs:21-214).
//
// The fixed schedules are the reference's: one two-sided rotor for 2x2,
// the 3-pair cycle for 3x3 and the three conflict-free layers for 4x4, up
// to 24 sweeps.  A matrix whose residual off-diagonal energy stays above
// 1e-10 * max(||A||_F, 1) escalates to LieSvdSmall (:90-92): it is appended
// to a fallback queue that a small_svd launch drains later in the same
// stream (or graph).  The sweep limit of n = 3, 4 can be lowered from the
// host (the handle's micro_max_sweeps) to force that escalation in tests.
#include <hip/hip_runtime.h>
#include <math.h>

#include "lsvd_kernels.hpp"

namespace lsvd {

__device__ __forceinline__ double wrap_half_pi(double a) {
    while (a > M_PI / 2) a -= M_PI;
    while (a < -M_PI / 2) a += M_PI;
    return a;
}

template <int N>
__device__ __forceinline__ double micro_offdiag(const double (&w)[16]) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j)
            if (i != j) s += w[i * 4 + j] * w[i * 4 + j];
    return sqrt(s);
}

// Left rotor on rows (i, j) and the U basis columns (:141-162); right rotor
// on columns (i, j) and the V basis columns (:164-185).
template <int N>
__device__ __forceinline__ void rotor_pair(double (&w)[16], double (&ub)[16], double (&vb)[16], int i, int j,
                                           double tl, double tr) {
    double sl, cl, sr, cr;
    sincos(tl, &sl, &cl);
#pragma unroll
    for (int c = 0; c < N; ++c) {
        const double ai = w[i * 4 + c], aj = w[j * 4 + c];
        w[i * 4 + c] = cl * ai - sl * aj;
        w[j * 4 + c] = sl * ai + cl * aj;
    }
#pragma unroll
    for (int r = 0; r < N; ++r) {
        const double ui = ub[r * 4 + i], uj = ub[r * 4 + j];
        ub[r * 4 + i] = cl * ui - sl * uj;
        ub[r * 4 + j] = sl * ui + cl * uj;
    }
    sincos(tr, &sr, &cr);
#pragma unroll
    for (int r = 0; r < N; ++r) {
        const double ai = w[r * 4 + i], aj = w[r * 4 + j];
        w[r * 4 + i] = cr * ai - sr * aj;
        w[r * 4 + j] = sr * ai + cr * aj;
    }
#pragma unroll
    for (int r = 0; r < N; ++r) {
        const double vi = vb[r * 4 + i], vj = vb[r * 4 + j];
        vb[r * 4 + i] = cr * vi - sr * vj;
        vb[r * 4 + j] = sr * vi + cr * vj;
    }
}

// Returns true when converged (false: escalate to LieSvdSmall).
template <int N>
__device__ bool micro_solve(const double* a, double* u, double* s, double* vt, int& rotations, int sweep_cap) {
    constexpr int kPairs = N == 2 ? 1 : (N == 3 ? 3 : 6);
    const int kSweeps = N == 2 ? 1 : sweep_cap;
    constexpr int pi_[6] = {0, 2, 0, 1, 0, 1};  // 4x4 layers {(0,1),(2,3)},{(0,2),(1,3)},{(0,3),(1,2)}
    constexpr int pj_[6] = {1, 3, 2, 3, 3, 2};
    constexpr int qi_[3] = {0, 0, 1};            // 3x3 cycle (0,1),(0,2),(1,2)
    constexpr int qj_[3] = {1, 2, 2};
    double w[16], ub[16], vb[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        w[i] = 0.0;
        ub[i] = (i % 5 == 0) ? 1.0 : 0.0;
        vb[i] = ub[i];
    }
    double fro = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) {
            w[i * 4 + j] = a[i * N + j];
            fro += w[i * 4 + j] * w[i * 4 + j];
        }
    const double ref = fmax(sqrt(fro), 1e-300);
    const double tol = kMicroTol * fmax(ref, 1.0);
    rotations = 0;
    for (int sw = 0; sw < kSweeps; ++sw) {
        const double energy = micro_offdiag<N>(w);
        if (energy <= tol || !isfinite(energy)) break;
        bool changed = false;
#pragma unroll
        for (int pp = 0; pp < kPairs; ++pp) {
            const int i = N == 4 ? pi_[pp] : (N == 3 ? qi_[pp] : 0);
            const int j = N == 4 ? pj_[pp] : (N == 3 ? qj_[pp] : 1);
            if (fabs(w[i * 4 + j]) + fabs(w[j * 4 + i]) <= tol) continue;
            // local_pair_svd_angles (:128-139)
            const double aa = w[i * 4 + i], bb = w[i * 4 + j], cc = w[j * 4 + i], dd = w[j * 4 + j];
            const double sum_angle = atan2(-(bb + cc), aa - dd);
            const double diff_angle = atan2(bb - cc, aa + dd);
            const double tl = wrap_half_pi(0.5 * (sum_angle + diff_angle));
            const double tr = wrap_half_pi(0.5 * (sum_angle - diff_angle));
            if (fabs(tl) + fabs(tr) <= 1e-18) continue;
            rotor_pair<N>(w, ub, vb, i, j, tl, tr);
            changed = true;
            ++rotations;
        }
        if (!changed) break;
    }
    if (micro_offdiag<N>(w) > 1e-10 * fmax(ref, 1.0)) return false;
    // extract_sorted_svd (:187-214): stable sort by |diag| descending.  The
    // rank of each diagonal entry is computed with constant indices only, so
    // w/ub/vb stay in registers; the dynamic index lands in global stores.
    double key[N];
#pragma unroll
    for (int i = 0; i < N; ++i) key[i] = fabs(w[i * 4 + i]);
#pragma unroll
    for (int src = 0; src < N; ++src) {
        int dst = 0;
#pragma unroll
        for (int j = 0; j < N; ++j) dst += (key[j] > key[src]) || (j < src && key[j] == key[src]);
        const double d = w[src * 4 + src];
        s[dst] = fabs(d);
        const double sign = d < 0.0 ? -1.0 : 1.0;
#pragma unroll
        for (int r = 0; r < N; ++r) {
            u[r * N + dst] = sign * ub[r * 4 + src];
            vt[dst * N + r] = vb[r * 4 + src];
        }
    }
    return true;
}

__global__ __launch_bounds__(256) void micro_svd_kernel(const Item* __restrict__ items, int count, Pools P,
                                                         Item* fallback, int* fallback_count, int sweep_cap) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= count) return;
    const Item itm = items[b];
    const int n = itm.n;
    const double* a = P.a + itm.a_off;
    double* u = P.u + itm.u_off;
    double* s = P.s + itm.s_off;
    double* vt = P.vt + itm.vt_off;
    int rot = 0;
    bool ok = true;
    switch (n) {
        case 1: {  // (:32-40)
            const double d = a[0];
            u[0] = d < 0.0 ? -1.0 : 1.0;
            s[0] = fabs(d);
            vt[0] = 1.0;
            break;
        }
        case 2:
            ok = micro_solve<2>(a, u, s, vt, rot, sweep_cap);
            break;
        case 3:
            ok = micro_solve<3>(a, u, s, vt, rot, sweep_cap);
            break;
        case 4:
            ok = micro_solve<4>(a, u, s, vt, rot, sweep_cap);
            break;
        default:
            break;
    }
    if (!ok) {
        Item f = itm;
        f.mode = kModeMicroFallback;
        const int slot = atomicAdd(fallback_count, 1);
        fallback[slot] = f;
    }
    if (P.info) {
        lsvd_info_t inf = {};
        inf.rotations = rot;
        inf.path = ok ? 1 : 2;
        inf.micro_fallback = ok ? 0 : 1;
        if (ok) P.info[itm.idx] = inf;  // a fallback's small solve writes its own entry
    }
}

hipError_t launch_micro(const MicroLaunch& L, hipStream_t st) {
    if (L.count <= 0) return hipSuccess;
    const int nt = 256;
    const int cap = L.sweep_cap >= 1 && L.sweep_cap <= 24 ? L.sweep_cap : 24;
    hipLaunchKernelGGL(micro_svd_kernel, dim3((L.count + nt - 1) / nt), dim3(nt), 0, st, L.items, L.count, L.pools,
                       L.fallback, L.fallback_count, cap);
    return hipGetLastError();
}

}  // namespace lsvd
