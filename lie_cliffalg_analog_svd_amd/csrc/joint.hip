// joint.hip -- Phase-JADE (LieSvdJoint::diagonalize_symmetric_with_params,
// src/lie_svd_joint.rs:71-158) for many families at once on gfx950.
//
// The reference visits the pairs of each round-robin layer one after the
// other.  The pairs of a layer are disjoint, and every quantity a pair's
// decision reads is untouched by the other pairs of its layer:
//   - the 2x2 blocks [M_ii M_ij; M_ji M_jj] of all K matrices (skip test,
//     Cardoso-Souloumiac angle, pi/4 alternative, trial rotations);
//   - the local axis energy of (i, j): rows i, j and columns i, j outside
//     the 2x2 block only rotate among themselves under any rotation of the
//     layer, so their squared sums are invariant, and the accept test
//     `after <= before + slack` reduces to a test on the 2x2 blocks.
// So each layer runs as two launches: `decide` (one wavefront per pair,
// lanes over the K matrices) picks every pair's accepted angle, and
// `apply` rotates the rows and columns of all K matrices and the basis,
// streaming each matrix through once and emitting the per-row
// off-diagonal energies the next layer's decisions and the sweep test need.
#include <hip/hip_runtime.h>
#include <math.h>

#include "lsvd_device.hpp"
#include "lsvd_joint.hpp"

namespace lsvd {

__device__ __forceinline__ double wrap_quarter_pi(double a) {  // wrap_jacobi_angle (:506-514)
    while (a > M_PI / 4) a -= M_PI / 2;
    while (a < -M_PI / 4) a += M_PI / 2;
    return a;
}

// layer_pairs (:524-539) in the reference's order; slot -> (i, j), i < j,
// or i = -1 when the slot pairs with the phantom index of odd n.
__device__ __forceinline__ void joint_pair(int n, int layer, int slot, int& i, int& j) {
    int p, q;
    rr_pair(n, layer, slot, p, q);
    if (q < 0) {
        i = -1;
        j = -1;
    } else {
        i = p;
        j = q;
    }
}

// ---- init: W = M, basis = I, per-row energies, per-matrix norms ---------
__global__ __launch_bounds__(256) void joint_init_kernel(JointArgs a) {
    const int f = blockIdx.y, k = blockIdx.x;
    const int n = a.n;
    const size_t nn = (size_t)n * n;
    const double* m = a.mats + ((size_t)f * a.K + k) * nn;
    double* w = a.work + ((size_t)f * a.K + k) * nn;
    __shared__ double red[8];
    double fro = 0.0, off = 0.0;
    for (int r = 0; r < n; ++r) {
        double ro = 0.0;
        for (int c = threadIdx.x; c < n; c += blockDim.x) {
            const double x = m[(size_t)r * n + c];
            w[(size_t)r * n + c] = x;
            fro += x * x;
            if (c != r) ro += x * x;
        }
        ro = block_sum<256>(ro, red);
        if (threadIdx.x == 0) a.offrow[((size_t)f * a.K + k) * n + r] = ro;
        off += ro;  // identical in every thread
    }
    fro = block_sum<256>(fro, red);
    if (threadIdx.x == 0) {
        a.part[((size_t)f * a.K + k) * 2 + 0] = fro;
        a.part[((size_t)f * a.K + k) * 2 + 1] = off;
    }
    if (k == 0) {
        double* v = a.basis + (size_t)f * nn;
        for (size_t e = threadIdx.x; e < nn; e += blockDim.x) v[e] = (e / n == e % n) ? 1.0 : 0.0;
    }
}

// ---- setup: reference norms and tolerances per family (:103-108) --------
__global__ void joint_setup_kernel(JointArgs a) {
    const int f = blockIdx.x;
    __shared__ double red[8];
    double fro = 0.0, off = 0.0;
    for (int k = threadIdx.x; k < a.K; k += blockDim.x) {
        fro += a.part[((size_t)f * a.K + k) * 2 + 0];
        off += a.part[((size_t)f * a.K + k) * 2 + 1];
    }
    fro = block_sum<64>(fro, red);
    off = block_sum<64>(off, red);
    if (threadIdx.x == 0) {
        JointFamily& s = a.fam[f];
        const double ref = fmax(sqrt(fro), 1e-300);
        const double scale = a.rel_tol ? ref : fmax(ref, 1.0);
        s.pair_tol = a.min_pair_scale * scale;
        s.tol = a.tol * scale;
        s.initial_offdiag = sqrt(off);
        s.before = sqrt(off);
        s.final_offdiag = sqrt(off);
        s.sweeps = 0;
        s.rotations = 0;
        s.rejected = 0;
        s.done = a.n <= 1 ? 1 : 0;
    }
}

// ---- decide: one wavefront per (family, pair) ----------------------------
__global__ __launch_bounds__(64) void joint_decide_kernel(JointArgs a, int layer) {
    const int f = blockIdx.y, slot = blockIdx.x;
    const int n = a.n, K = a.K;
    const size_t nn = (size_t)n * n;
    JointRot* rot = a.rot + (size_t)f * a.slots + slot;
    int i, j;
    joint_pair(n, layer, slot, i, j);
    if (a.fam[f].done || i < 0) {
        if (threadIdx.x == 0) *rot = JointRot{1.0, 0.0, 0};
        return;
    }
    const int lane = threadIdx.x;
    // gather the 2x2 blocks (lanes over k) and the pair's axis energies
    double sxx = 0, syy = 0, sxy = 0, poff = 0, loc = 0, old2 = 0;
    for (int k = lane; k < K; k += kWave) {
        const double* w = a.work + ((size_t)f * K + k) * nn;
        const double aa = w[(size_t)i * n + i], bb = w[(size_t)i * n + j];
        const double cc = w[(size_t)j * n + i], dd = w[(size_t)j * n + j];
        const double x = 0.5 * (aa - dd), y = 0.5 * (bb + cc);
        sxx += x * x;
        syy += y * y;
        sxy += x * y;
        poff += fabs(bb) + fabs(cc);
        const double* orow = a.offrow + ((size_t)f * K + k) * n;
        // local_offdiag_sq_for_axes (:475-494): rows i, j off the diagonal
        // plus columns i, j outside rows {i, j}; columns mirror rows.
        loc += 2.0 * orow[i] + 2.0 * orow[j] - bb * bb - cc * cc;
        old2 += bb * bb + cc * cc;
    }
    sxx = wave_sum(sxx);
    syy = wave_sum(syy);
    sxy = wave_sum(sxy);
    poff = wave_sum(poff);
    loc = wave_sum(loc);
    old2 = wave_sum(old2);
    const JointFamily& fs = a.fam[f];
    if (poff <= fs.pair_tol) {  // (:118)
        if (lane == 0) *rot = JointRot{1.0, 0.0, 0};
        return;
    }
    // joint_symmetric_pair_angle (:338-357)
    double theta = 0.25 * atan2(2.0 * sxy, syy - sxx);
    const double alt = wrap_quarter_pi(theta + M_PI / 4);
    theta = wrap_quarter_pi(theta);
    double e_alt = 0.0, e_th = 0.0;
    {
        double sa, ca, st, ct;
        sincos(alt, &sa, &ca);
        sincos(theta, &st, &ct);
        for (int k = lane; k < K; k += kWave) {
            const double* w = a.work + ((size_t)f * K + k) * nn;
            const double aa = w[(size_t)i * n + i], dd = w[(size_t)j * n + j];
            const double b = 0.5 * (w[(size_t)i * n + j] + w[(size_t)j * n + i]);
            const double o1 = 0.5 * (aa - dd) * (2.0 * sa * ca) + b * (ca * ca - sa * sa);
            const double o2 = 0.5 * (aa - dd) * (2.0 * st * ct) + b * (ct * ct - st * st);
            e_alt += 2.0 * o1 * o1;
            e_th += 2.0 * o2 * o2;
        }
        e_alt = wave_sum(e_alt);
        e_th = wave_sum(e_th);
    }
    if (e_alt < e_th) theta = alt;
    if (fabs(theta) <= 1e-18) {  // (:122)
        if (lane == 0) *rot = JointRot{1.0, 0.0, 0};
        return;
    }
    // accept_joint_rotor (:359-386): halve the angle until the local axis
    // energy does not grow (beyond slack); after - before = new2 - old2.
    const double before = loc;
    const double slack = 1e-14 * fmax(before, 1.0);
    double scale = 1.0;
    const int steps = a.line_search_steps > 1 ? a.line_search_steps : 1;
    for (int t = 0; t < steps; ++t) {
        const double ang = theta * scale;
        double s, c;
        sincos(ang, &s, &c);
        double new2 = 0.0;
        for (int k = lane; k < K; k += kWave) {
            const double* w = a.work + ((size_t)f * K + k) * nn;
            const double aa = w[(size_t)i * n + i], bb = w[(size_t)i * n + j];
            const double cc = w[(size_t)j * n + i], dd = w[(size_t)j * n + j];
            // left rotor on rows (i, j), then right rotor on columns (i, j)
            const double l00 = c * aa - s * cc, l01 = c * bb - s * dd;
            const double l10 = s * aa + c * cc, l11 = s * bb + c * dd;
            const double r01 = s * l00 + c * l01;
            const double r10 = c * l10 - s * l11;
            new2 += r01 * r01 + r10 * r10;
        }
        new2 = wave_sum(new2);
        const double after = before - old2 + new2;
        if (new2 - old2 <= slack && isfinite(after)) {
            if (lane == 0) {
                *rot = JointRot{c, s, 1};
                atomicAdd(&a.fam[f].rotations, 1);
            }
            return;
        }
        scale *= 0.5;
    }
    // rejected (:383-385): the reference rotated rows and columns (i, j) and
    // rotated them back.  In finite arithmetic that restores them up to
    // rounding, and the accept test only fails on non-finite energies, so
    // what the round trip leaves is its NaN pattern: a NaN angle turns
    // rows and columns i, j of every matrix into NaN (flag 2 with NaN c, s);
    // a finite angle mixes NaNs between the two rows and the two columns,
    // which the NaN-propagating identity (c, s) = (1, 0) does (0 * NaN =
    // NaN, 0 * x = 0).  The basis is left alone (:377-379).
    if (lane == 0) {
        const double nan = __builtin_nan("");
        *rot = isfinite(theta) ? JointRot{1.0, 0.0, 2} : JointRot{nan, nan, 2};
        atomicAdd(&a.fam[f].rejected, 1);
    }
}

// ---- apply: rotate SPB row pairs of one matrix (or of the basis) ----------
// Row pair (i, j) of slot s: left rotor on the two rows, then the right
// rotor of every column's own pair (exchanged through LDS), then per-row
// off-diagonal energies.  blockIdx.z == K selects the basis, which takes
// the right rotor only (apply_basis_rotor, :428-437).
template <int SPB>
__global__ __launch_bounds__(256) void joint_apply_kernel(JointArgs a, int layer) {
    extern __shared__ double rows[];  // [2*SPB][n]
    const int f = blockIdx.y, k = blockIdx.z, chunk = blockIdx.x;
    const int n = a.n, K = a.K;
    if (a.fam[f].done) return;
    const size_t nn = (size_t)n * n;
    const bool is_basis = k == K;
    double* w = is_basis ? a.basis + (size_t)f * nn : a.work + ((size_t)f * K + k) * nn;
    const JointRot* rt = a.rot + (size_t)f * a.slots;
    // rows handled by this block: basis -> plain row ranges; matrices -> slot pairs
    const int s0 = chunk * SPB;
    int rid[2 * SPB];
#pragma unroll
    for (int q = 0; q < SPB; ++q) {
        int i = -1, j = -1;
        if (is_basis) {
            i = 2 * (s0 + q) < n ? 2 * (s0 + q) : -1;
            j = 2 * (s0 + q) + 1 < n ? 2 * (s0 + q) + 1 : -1;
        } else if (s0 + q < a.slots) {
            int pp, qq;
            rr_pair(n, layer, s0 + q, pp, qq);  // qq == -1: unpaired row of odd n
            i = pp;
            j = qq;
        }
        rid[2 * q] = i;
        rid[2 * q + 1] = j;
    }
    // 1. load + left rotor (matrices only)
#pragma unroll
    for (int q = 0; q < SPB; ++q) {
        const int i = rid[2 * q], j = rid[2 * q + 1];
        double c = 1.0, s = 0.0;
        bool rotate = false;  // flag 0 (skipped): rows copied unchanged
        if (!is_basis && s0 + q < a.slots && j >= 0) {
            const JointRot r = rt[s0 + q];
            c = r.c;
            s = r.s;
            rotate = r.flag != 0;
        }
        for (int col = threadIdx.x; col < n; col += blockDim.x) {
            const double xi = i >= 0 ? w[(size_t)i * n + col] : 0.0;
            const double xj = j >= 0 ? w[(size_t)j * n + col] : 0.0;
            rows[(2 * q) * n + col] = rotate ? c * xi - s * xj : xi;
            rows[(2 * q + 1) * n + col] = rotate ? s * xi + c * xj : xj;
        }
    }
    __syncthreads();
    // 2. right rotor per column pair of this layer, write back, energies
    double ro[2 * SPB];
#pragma unroll
    for (int e = 0; e < 2 * SPB; ++e) ro[e] = 0.0;
    for (int col = threadIdx.x; col < n; col += blockDim.x) {
        // column pair of `col` in this layer
        int partner = -1, is_first = 1;
        double c = 1.0, s = 0.0;
        {
            // invert the circle-method placement: which slot holds `col`
            const int m = n + (n & 1), ring = m - 1;
            int slot, other;
            if (col == m - 1) {
                slot = 0;
                other = layer % ring;
            } else {
                const int rel = ((col - layer) % ring + ring) % ring;  // col = (layer + rel) % ring
                if (rel == 0) {
                    slot = 0;
                    other = m - 1;
                } else if (rel <= m / 2 - 1) {  // col is the first index of slot rel
                    slot = rel;
                    other = (layer + ring - rel) % ring;
                } else {  // col is the second index of slot ring - rel
                    slot = ring - rel;
                    other = (layer + slot) % ring;
                }
            }
            if (other < n) {
                const JointRot r = rt[slot];
                // matrices take accepted and rejected rotors (flag 1, 2), the
                // basis accepted ones only
                if (r.flag == 1 || (r.flag == 2 && !is_basis)) {
                    partner = other;
                    is_first = col < other;
                    c = r.c;
                    s = r.s;
                }
            }
        }
#pragma unroll
        for (int e = 0; e < 2 * SPB; ++e) {
            const int row = rid[e];
            if (row < 0) continue;
            const double x = rows[e * n + col];
            double y = x;
            if (partner >= 0) {
                const double xp = rows[e * n + partner];
                // apply_right_rotor: col_i = c col_i - s col_j; col_j = s col_i + c col_j
                y = is_first ? c * x - s * xp : s * xp + c * x;
            }
            w[(size_t)row * n + col] = y;
            if (col != row) ro[e] += y * y;
        }
    }
    if (!is_basis) {
        // row energies: wave sums, then one pass over the four waves' partials
        __shared__ double part[2 * SPB][4];
        const int wv = threadIdx.x >> 6;
#pragma unroll
        for (int e = 0; e < 2 * SPB; ++e) {
            const double t = wave_sum(ro[e]);
            if ((threadIdx.x & 63) == 0) part[e][wv] = t;
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < 2 * SPB; ++e)
            if (threadIdx.x == e && rid[e] >= 0)
                a.offrow[((size_t)f * K + k) * n + rid[e]] = part[e][0] + part[e][1] + part[e][2] + part[e][3];
    }
}

// ---- sweep end: off-diagonal norm and the stop test (:139-142) ----------
__global__ void joint_sweep_end_kernel(JointArgs a) {
    const int f = blockIdx.x;
    __shared__ double red[8];
    if (a.fam[f].done) return;
    double off = 0.0;
    const size_t cnt = (size_t)a.K * a.n;
    for (size_t e = threadIdx.x; e < cnt; e += blockDim.x) off += a.offrow[(size_t)f * cnt + e];
    off = block_sum<256>(off, red);
    if (threadIdx.x == 0) {
        JointFamily& s = a.fam[f];
        const double after = sqrt(off);
        s.sweeps += 1;
        s.final_offdiag = after;
        if (after <= s.tol || after >= s.before * (1.0 - 1e-10) || s.sweeps >= a.max_sweeps) s.done = 1;
        s.before = after;
    }
}

// ---- final: diagonals --------------------------------------------------------
__global__ void joint_diag_kernel(JointArgs a, double* diag) {
    const int f = blockIdx.y;
    const size_t nn = (size_t)a.n * a.n;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < a.K * a.n; e += gridDim.x * blockDim.x) {
        const int k = e / a.n, i = e % a.n;
        diag[((size_t)f * a.K + k) * a.n + i] = a.work[((size_t)f * a.K + k) * nn + (size_t)i * a.n + i];
    }
}

// ---- the layer-pair map itself (diagnostic) -----------------------------------
__global__ void joint_pairs_kernel(int n, int layers, int slots, int32_t* out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= layers * slots) return;
    int i, j;
    joint_pair(n, t / slots, t % slots, i, j);
    out[2 * t] = i;
    out[2 * t + 1] = j;
}

hipError_t joint_launch_pairs(int n, int32_t* out, hipStream_t st) {
    const int layers = n % 2 == 0 ? n - 1 : n, slots = (n + 1) / 2;
    const int total = layers * slots;
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(joint_pairs_kernel, dim3((total + 255) / 256), dim3(256), 0, st, n, layers, slots, out);
    return hipGetLastError();
}

// ---- launches --------------------------------------------------------------------
constexpr int kSPB = 2;

hipError_t joint_launch_init(const JointArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(joint_init_kernel, dim3(a.K, a.F), dim3(256), 0, st, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(joint_setup_kernel, dim3(a.F), dim3(64), 0, st, a);
    return hipGetLastError();
}

int joint_layers(const JointArgs& a) { return a.n % 2 == 0 ? a.n - 1 : a.n; }

hipError_t joint_launch_layers(const JointArgs& a, hipStream_t st, int first, int count, bool sweep_end) {
    const int chunks = (a.slots + kSPB - 1) / kSPB;
    const int basis_chunks = ((a.n + 1) / 2 + kSPB - 1) / kSPB;
    const int cz = chunks > basis_chunks ? chunks : basis_chunks;
    const size_t lds = sizeof(double) * 2 * kSPB * a.n;
    for (int layer = first; layer < first + count; ++layer) {
        hipLaunchKernelGGL(joint_decide_kernel, dim3(a.slots, a.F), dim3(64), 0, st, a, layer);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(joint_apply_kernel<kSPB>, dim3(cz, a.F, a.K + 1), dim3(256), lds, st, a, layer);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (!sweep_end) return hipSuccess;
    hipLaunchKernelGGL(joint_sweep_end_kernel, dim3(a.F), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t joint_launch_diag(const JointArgs& a, double* diag, hipStream_t st) {
    hipLaunchKernelGGL(joint_diag_kernel, dim3(((a.K * a.n) + 255) / 256, a.F), dim3(256), 0, st, a, diag);
    return hipGetLastError();
}

}  // namespace lsvd
