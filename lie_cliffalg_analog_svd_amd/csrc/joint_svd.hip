// joint_svd.hip -- two-sided family SVD (LieSvdJoint::joint_svd_with_params,
// src/lie_svd_joint.rs:170-336) for many families at once on gfx950.
//
// A family is K matrices of rows x cols; the sweeps look for a shared U
// (rows x rows) and V (cols x cols) with U^T M_k V as diagonal as possible,
// pairing indices below kk = min(rows, cols) on the round-robin layers.  As
// for the symmetric Phase-JADE (joint.hip), the pairs of a layer are
// disjoint and a pair's decision only reads quantities the other pairs of
// its layer leave unchanged:
//   - the 2x2 blocks [M_ii M_ij; M_ji M_jj] (skip test, weighted averaged
//     block, two-sided angles, trial rotations);
//   - its local axis energy: entries of rows i, j outside columns {i, j}
//     rotate among themselves under the left rotor, entries of columns i, j
//     outside rows {i, j} under the right rotor, so only the block's
//     off-diagonal pair changes and `after <= before + slack` reduces to
//     new(M_ij^2 + M_ji^2) - old(M_ij^2 + M_ji^2) <= slack, with `before`
//     assembled from per-row and per-column off-diagonal energies.
// So a layer is `decide` (one wavefront per pair, lanes over K) and `apply`
// (one wavefront per row group of a matrix or basis -- a row pair, or two
// rows that take no left rotor -- with its lanes over the layer's column
// groups: every 2x2 block rotated in place, the two rows' off-diagonal
// energies by wave sums, and column sums of squares per workgroup, which
// `decide` adds up over the row blocks).
#include <hip/hip_runtime.h>
#include <math.h>

#include "lsvd_device.hpp"
#include "lsvd_joint_svd.hpp"

namespace lsvd {

__device__ __forceinline__ double wrap_half_pi_d(double a) {  // wrap_svd_angle (:291-299)
    while (a > M_PI / 2) a -= M_PI;
    while (a < -M_PI / 2) a += M_PI;
    return a;
}

// Layer position of index g < kk: its partner at this layer (-1 if
// unpaired) -- the inverse of the circle-method placement.
__device__ __forceinline__ int jsvd_partner(int kk, int layer, int g) {
    const int m = kk + (kk & 1), ring = m - 1;
    int other;
    if (g == m - 1) {
        other = layer % ring;
    } else {
        const int rel = ((g - layer) % ring + ring) % ring;
        if (rel == 0) other = m - 1;
        else if (rel <= m / 2 - 1) other = (layer + ring - rel) % ring;
        else other = (layer + (ring - rel)) % ring;
    }
    return other < kk ? other : -1;
}

// Slot of the pair holding g (valid when g is paired).
__device__ __forceinline__ int jsvd_slot(int kk, int layer, int g) {
    const int m = kk + (kk & 1), ring = m - 1;
    if (g == m - 1) return 0;
    const int rel = ((g - layer) % ring + ring) % ring;
    if (rel == 0) return 0;
    return rel <= m / 2 - 1 ? rel : ring - rel;
}

// ---- init: W = M, U = I, V = I, row/column energies, norms ---------------
__global__ __launch_bounds__(256) void jsvd_init_kernel(JsvdArgs a) {
    const int f = blockIdx.y, k = blockIdx.x;
    const int R = a.rows, Cc = a.cols;
    const size_t rc = (size_t)R * Cc;
    const double* m = a.mats + ((size_t)f * a.K + k) * rc;
    double* w = a.work + ((size_t)f * a.K + k) * rc;
    __shared__ double red[8];
    double fro = 0.0, off = 0.0;
    for (size_t e = threadIdx.x; e < rc; e += blockDim.x) {
        const double x = m[e];
        w[e] = x;
        fro += x * x;
        if ((int)(e / Cc) != (int)(e % Cc)) off += x * x;
    }
    fro = block_sum<256>(fro, red);
    off = block_sum<256>(off, red);
    if (threadIdx.x == 0) {
        a.part[((size_t)f * a.K + k) * 2 + 0] = fro;
        a.part[((size_t)f * a.K + k) * 2 + 1] = off;
    }
    // row and column off-diagonal energies
    double* ro = a.rowoff + ((size_t)f * a.K + k) * R;
    double* co = a.colpart + ((size_t)f * a.K + k) * a.rb * Cc;
    for (int r = threadIdx.x; r < R; r += blockDim.x) {
        double s = 0.0;
        for (int c = 0; c < Cc; ++c)
            if (c != r) s += m[(size_t)r * Cc + c] * m[(size_t)r * Cc + c];
        ro[r] = s;
    }
    for (int c = threadIdx.x; c < Cc; c += blockDim.x) {
        double s = 0.0;
        for (int r = 0; r < R; ++r) s += m[(size_t)r * Cc + c] * m[(size_t)r * Cc + c];
        co[c] = s;
        for (int b = 1; b < a.rb; ++b) co[(size_t)b * Cc + c] = 0.0;
    }
    if (k == 0) {
        double* u = a.ubasis + (size_t)f * R * R;
        double* v = a.vbasis + (size_t)f * Cc * Cc;
        for (size_t e = threadIdx.x; e < (size_t)R * R; e += blockDim.x) u[e] = (e / R == e % R) ? 1.0 : 0.0;
        for (size_t e = threadIdx.x; e < (size_t)Cc * Cc; e += blockDim.x) v[e] = (e / Cc == e % Cc) ? 1.0 : 0.0;
    }
}

// ---- setup: reference norms and tolerances per family (:201-206) --------
__global__ void jsvd_setup_kernel(JsvdArgs a) {
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
        s.pair_tol = a.min_pair_scale * fmax(ref, 1.0);
        s.tol = a.tol * fmax(ref, 1.0);
        s.initial_offdiag = sqrt(off);
        s.before = sqrt(off);
        s.final_offdiag = sqrt(off);
        s.sweeps = 0;
        s.rotations = 0;
        s.rejected = 0;
        s.done = a.kk <= 1 ? 1 : 0;
    }
}

// ---- decide: one wavefront per (family, pair) ----------------------------
__global__ __launch_bounds__(64) void jsvd_decide_kernel(JsvdArgs a, int layer) {
    const int f = blockIdx.y, slot = blockIdx.x;
    const int R = a.rows, Cc = a.cols, K = a.K;
    const size_t rc = (size_t)R * Cc;
    JsvdRot* rot = a.rot + (size_t)f * a.slots + slot;
    int i, j;
    rr_pair(a.kk, layer, slot, i, j);
    if (a.fam[f].done || j < 0) {
        if (threadIdx.x == 0) *rot = JsvdRot{1.0, 0.0, 1.0, 0.0, 0};
        return;
    }
    const int lane = threadIdx.x;
    // joint_pair_offdiag (:451-453), the weighted averaged block
    // (joint_svd_pair_angles :264-275) and local_offdiag_sq_for_axes
    // (:475-494) from the row/column energies
    double sa = 0, sb = 0, sc = 0, sd = 0, poff = 0, loc = 0, old2 = 0;
    for (int k = lane; k < K; k += kWave) {
        const double* w = a.work + ((size_t)f * K + k) * rc;
        const double aa = w[(size_t)i * Cc + i], bb = w[(size_t)i * Cc + j];
        const double cc = w[(size_t)j * Cc + i], dd = w[(size_t)j * Cc + j];
        poff += fabs(bb) + fabs(cc);
        const double wt = fmax(sqrt(aa * aa + bb * bb + cc * cc + dd * dd), 1e-300);  // pair_block_norm
        sa += aa / wt;
        sb += bb / wt;
        sc += cc / wt;
        sd += dd / wt;
        const double* ro = a.rowoff + ((size_t)f * K + k) * R;
        // coloff_i = S_i - M_ii^2 with S the column sums of squares
        loc += ro[i] + ro[j] - (aa * aa + cc * cc) - (dd * dd + bb * bb);
        old2 += bb * bb + cc * cc;
    }
    for (int t = lane; t < K * a.rb; t += kWave) {
        const double* co = a.colpart + ((size_t)f * K * a.rb + t) * Cc;
        loc += co[i] + co[j];
    }
    poff = wave_sum(poff);
    sa = wave_sum(sa);
    sb = wave_sum(sb);
    sc = wave_sum(sc);
    sd = wave_sum(sd);
    loc = wave_sum(loc);
    old2 = wave_sum(old2);
    const JointFamily& fs = a.fam[f];
    if (poff <= fs.pair_tol) {  // (:219-221)
        if (lane == 0) *rot = JsvdRot{1.0, 0.0, 1.0, 0.0, 0};
        return;
    }
    // local_pair_svd_angles_from_values (:282-289)
    const double sum_angle = atan2(-(sb + sc), sa - sd);
    const double diff_angle = atan2(sb - sc, sa + sd);
    const double tl = wrap_half_pi_d(0.5 * (sum_angle + diff_angle));
    const double tr = wrap_half_pi_d(0.5 * (sum_angle - diff_angle));
    if (fabs(tl) + fabs(tr) <= 1e-18) {  // (:222-224)
        if (lane == 0) *rot = JsvdRot{1.0, 0.0, 1.0, 0.0, 0};
        return;
    }
    // accept_joint_svd_rotor (:303-336): halve both angles until the local
    // axis energy does not grow beyond the slack
    const double before = loc;
    const double slack = 1e-14 * fmax(before, 1.0);
    double scale = 1.0;
    const int steps = a.line_search_steps > 1 ? a.line_search_steps : 1;
    for (int t = 0; t < steps; ++t) {
        double sl, cl, sr, cr;
        sincos(tl * scale, &sl, &cl);
        sincos(tr * scale, &sr, &cr);
        double new2 = 0.0;
        for (int k = lane; k < K; k += kWave) {
            const double* w = a.work + ((size_t)f * K + k) * rc;
            const double aa = w[(size_t)i * Cc + i], bb = w[(size_t)i * Cc + j];
            const double cc = w[(size_t)j * Cc + i], dd = w[(size_t)j * Cc + j];
            // apply_left_rotor on rows (i, j), then apply_right_rotor on columns (i, j)
            const double l00 = cl * aa - sl * cc, l01 = cl * bb - sl * dd;
            const double l10 = sl * aa + cl * cc, l11 = sl * bb + cl * dd;
            const double r01 = sr * l00 + cr * l01;
            const double r10 = cr * l10 - sr * l11;
            new2 += r01 * r01 + r10 * r10;
        }
        new2 = wave_sum(new2);
        const double after = before - old2 + new2;
        if (new2 - old2 <= slack && isfinite(after)) {
            if (lane == 0) {
                *rot = JsvdRot{cl, sl, cr, sr, 1};
                atomicAdd(&a.fam[f].rotations, 1);
            }
            return;
        }
        scale *= 0.5;
    }
    if (lane == 0) {
        *rot = JsvdRot{1.0, 0.0, 1.0, 0.0, 0};
        atomicAdd(&a.fam[f].rejected, 1);
    }
}

// ---- apply: blockIdx.x < K rb: matrix blockIdx.x / rb, row block
// blockIdx.x % rb; then the U basis' rbu blocks, then the V basis' rbv.
// One wavefront per row task (see jsvd_matrix_row_tasks); lanes over the
// column groups: the layer's slots (a pair, or the odd index alone), then
// the columns >= kk alone.  U's columns take the left angle, V's and the
// matrices' columns the right angle; only the matrices' rows take the left
// angle.  Every block is read, rotated and written by one lane, in place.
__global__ __launch_bounds__(256) void jsvd_apply_kernel(JsvdArgs a, int layer) {
    extern __shared__ double colp[];  // kJsvdTasks x cols column sums
    const int f = blockIdx.y;
    if (a.fam[f].done) return;
    const int K = a.K, kk = a.kk;
    int bx = blockIdx.x;
    int kind, x = 0;  // 0 matrix x, 1 U, 2 V
    if (bx < K * a.rb) {
        kind = 0;
        x = bx / a.rb;
        bx -= x * a.rb;
    } else if (bx < K * a.rb + a.rbu) {
        kind = 1;
        bx -= K * a.rb;
    } else {
        kind = 2;
        bx -= K * a.rb + a.rbu;
    }
    const int R = kind == 2 ? a.cols : a.rows;
    const int Cc = kind == 1 ? a.rows : a.cols;
    double* w = kind == 1 ? a.ubasis + (size_t)f * a.rows * a.rows
                          : (kind == 2 ? a.vbasis + (size_t)f * a.cols * a.cols
                                       : a.work + ((size_t)f * K + x) * a.rows * a.cols);
    const JsvdRot* rt = a.rot + (size_t)f * a.slots;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int task = bx * kJsvdTasks + wv;
    // the wave's rows r0, r1 (-1: none) and their left rotor
    int r0 = -1, r1 = -1;
    double cl = 1.0, sl = 0.0;
    if (kind == 0 && task < a.slots) {
        int p, q;
        rr_pair(kk, layer, task, p, q);
        r0 = p;
        if (q >= 0) {
            r1 = q;
            const JsvdRot t = rt[task];
            cl = t.cl;
            sl = t.sl;
        }
    } else {
        const int base = kind == 0 ? kk + 2 * (task - a.slots) : 2 * task;
        if (base < R) r0 = base;
        if (base + 1 < R) r1 = base + 1;
    }
    double e0 = 0.0, e1 = 0.0;  // off-diagonal energies of rows r0, r1
    const int ctasks = a.slots + (Cc - kk);
    for (int t = lane; t < ctasks; t += kWave) {
        int c0, c1 = -1;
        double cr = 1.0, sr = 0.0;
        if (t < a.slots) {
            int p, q;
            rr_pair(kk, layer, t, p, q);
            c0 = p;
            if (q >= 0) {
                c1 = q;
                const JsvdRot g = rt[t];
                cr = kind == 1 ? g.cl : g.cr;
                sr = kind == 1 ? g.sl : g.sr;
            }
        } else {
            c0 = kk + (t - a.slots);
        }
        const bool h0 = r0 >= 0, h1 = r1 >= 0, k1 = c1 >= 0;
        const double b00 = h0 ? w[(size_t)r0 * Cc + c0] : 0.0;
        const double b01 = (h0 && k1) ? w[(size_t)r0 * Cc + c1] : 0.0;
        const double b10 = h1 ? w[(size_t)r1 * Cc + c0] : 0.0;
        const double b11 = (h1 && k1) ? w[(size_t)r1 * Cc + c1] : 0.0;
        // left: row_i = c row_i - s row_j, row_j = s row_i + c row_j
        const double l00 = cl * b00 - sl * b10, l01 = cl * b01 - sl * b11;
        const double l10 = sl * b00 + cl * b10, l11 = sl * b01 + cl * b11;
        // right: col_i = c col_i - s col_j, col_j = s col_i + c col_j
        const double y00 = cr * l00 - sr * l01, y01 = sr * l00 + cr * l01;
        const double y10 = cr * l10 - sr * l11, y11 = sr * l10 + cr * l11;
        if (h0) w[(size_t)r0 * Cc + c0] = y00;
        if (h0 && k1) w[(size_t)r0 * Cc + c1] = y01;
        if (h1) w[(size_t)r1 * Cc + c0] = y10;
        if (h1 && k1) w[(size_t)r1 * Cc + c1] = y11;
        if (kind == 0) {
            e0 += (c0 != r0 ? y00 * y00 : 0.0) + (k1 && c1 != r0 ? y01 * y01 : 0.0);
            e1 += (c0 != r1 ? y10 * y10 : 0.0) + (k1 && c1 != r1 ? y11 * y11 : 0.0);
            colp[wv * Cc + c0] = y00 * y00 + y10 * y10;
            if (k1) colp[wv * Cc + c1] = y01 * y01 + y11 * y11;
        }
    }
    if (kind != 0) return;
    e0 = wave_sum(e0);
    e1 = wave_sum(e1);
    double* ro = a.rowoff + ((size_t)f * K + x) * a.rows;
    if (lane == 0) {
        if (r0 >= 0) ro[r0] = e0;
        if (r1 >= 0) ro[r1] = e1;
    }
    __syncthreads();
    double* co = a.colpart + (((size_t)f * K + x) * a.rb + bx) * Cc;
    for (int c = threadIdx.x; c < Cc; c += blockDim.x) {
        double s = 0.0;
#pragma unroll
        for (int v = 0; v < kJsvdTasks; ++v) s += colp[v * Cc + c];
        co[c] = s;
    }
}

// ---- sweep end: off-diagonal norm and the stop test (:246-249) ----------
__global__ void jsvd_sweep_end_kernel(JsvdArgs a) {
    const int f = blockIdx.x;
    __shared__ double red[8];
    if (a.fam[f].done) return;
    double off = 0.0;
    const size_t cnt = (size_t)a.K * a.rows;
    for (size_t e = threadIdx.x; e < cnt; e += blockDim.x) off += a.rowoff[(size_t)f * cnt + e];
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

// ---- final: |diagonal| (diagonal_corridor :501-504), V^T ------------------
__global__ void jsvd_final_kernel(JsvdArgs a, double* sig, double* vt) {
    const int f = blockIdx.y;
    const size_t rc = (size_t)a.rows * a.cols;
    const int tot = a.K * a.kk;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += gridDim.x * blockDim.x) {
        const int k = e / a.kk, i = e % a.kk;
        sig[(size_t)f * tot + e] = fabs(a.work[((size_t)f * a.K + k) * rc + (size_t)i * a.cols + i]);
    }
    const size_t cc = (size_t)a.cols * a.cols;
    const double* v = a.vbasis + (size_t)f * cc;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < cc; e += (size_t)gridDim.x * blockDim.x) {
        const size_t r = e / a.cols, c = e % a.cols;
        vt[(size_t)f * cc + e] = v[c * a.cols + r];
    }
}

// ---- launches ---------------------------------------------------------------
hipError_t jsvd_launch_init(const JsvdArgs& a, hipStream_t st) {
    const size_t lds = sizeof(double) * kJsvdTasks * a.cols;  // the apply launches of the sweep graph
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&jsvd_apply_kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(jsvd_init_kernel, dim3(a.K, a.F), dim3(256), 0, st, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(jsvd_setup_kernel, dim3(a.F), dim3(64), 0, st, a);
    return hipGetLastError();
}

int jsvd_layers(const JsvdArgs& a) { return (a.kk & 1) ? a.kk : a.kk - 1; }

hipError_t jsvd_launch_layers(const JsvdArgs& a, hipStream_t st, int first, int count, bool sweep_end) {
    for (int layer = first; layer < first + count; ++layer) {
        hipLaunchKernelGGL(jsvd_decide_kernel, dim3(a.slots, a.F), dim3(64), 0, st, a, layer);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(jsvd_apply_kernel, dim3(a.K * a.rb + a.rbu + a.rbv, a.F), dim3(64 * kJsvdTasks),
                           sizeof(double) * kJsvdTasks * a.cols, st, a, layer);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (!sweep_end) return hipSuccess;
    hipLaunchKernelGGL(jsvd_sweep_end_kernel, dim3(a.F), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t jsvd_launch_final(const JsvdArgs& a, double* sig, double* vt, hipStream_t st) {
    const int work = std::max(a.K * a.kk, a.cols * a.cols);
    hipLaunchKernelGGL(jsvd_final_kernel, dim3((work + 255) / 256, a.F), dim3(256), 0, st, a, sig, vt);
    return hipGetLastError();
}

}  // namespace lsvd
