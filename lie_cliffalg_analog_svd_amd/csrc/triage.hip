// triage.hip -- the adaptive router's matrix triage (analyze_matrix,
// src/lie_svd_adaptive.rs:186-274, with phase_signature and
// global_phase_invariants, src/lie_svd_phasehealth.rs:100-345, and the Block4
// dual balance, src/lie_svd_block4.rs:210-262) for a batch of square
// matrices, with choose_route's decision on the result.
//
// One 256-thread workgroup per matrix.  Every statistic is a reduction over
// rows, columns, entries or 4 x 4 diagonal blocks: a wavefront takes one row
// (then one column) at a time for the per-vector phase health (mean, centred
// energy, one-step delay product, gradient, energy entropy, twist), keeping
// the row / column L1 and L2 norms in LDS; the whole workgroup then sweeps
// the entries (total, diagonal, off-diagonal, symmetric / skew, upper / lower
// energies, the global phase entropy) and the 4 x 4 blocks, and finally the
// norm vectors (coefficients of variation, L1 mismatch, mass entropies).
#include <hip/hip_runtime.h>
#include <math.h>

#include "lsvd_adaptive.hpp"
#include "lsvd_device.hpp"

namespace lsvd {
namespace {

constexpr int kNT = 256;
constexpr int kWaves = kNT / kWave;

struct VecStats {
    double l1, l2, delay_bivector_sq, gradient_sq, entropy, twist;
};

// analyze_vector_phase_by + normalized_energy_entropy_by
// (src/lie_svd_phasehealth.rs:146-211, 313-329) for x[k * stride], k < len,
// by one wavefront; every lane returns the same values.
__device__ VecStats wave_vector_phase(const double* x, int len, int stride, int lane) {
    VecStats s{};
    if (len == 0) return s;
    double sum = 0.0;
    for (int k = lane; k < len; k += kWave) sum += x[(int64_t)k * stride];
    const double mean = wave_sum(sum) / (double)len;
    double vsq = 0.0, nsq = 0.0, dly = 0.0, grd = 0.0, l1 = 0.0;
    for (int k = lane; k < len; k += kWave) {
        const double xi = x[(int64_t)k * stride];
        const double xn = x[(int64_t)((k + 1) % len) * stride];
        const double c = xi - mean;
        vsq += c * c;
        nsq += xi * xi;
        dly += xi * xn;
        const double d = xn - xi;
        grd += d * d;
        l1 += fabs(xi);
    }
    vsq = wave_sum(vsq);
    nsq = wave_sum(nsq);
    dly = wave_sum(dly);
    grd = wave_sum(grd);
    l1 = wave_sum(l1);
    double ent = 0.0;
    if (len > 1 && nsq > 1e-300) {
        for (int k = lane; k < len; k += kWave) {
            const double v = x[(int64_t)k * stride];
            const double p = (v * v) / nsq;
            if (p > 0.0) ent -= p * log(p);
        }
        ent = wave_sum(ent) / log((double)len);
    }
    s.l1 = l1;
    s.l2 = sqrt(nsq);
    s.delay_bivector_sq = fmax(nsq * nsq - dly * dly, 0.0);
    s.gradient_sq = grd;
    s.entropy = ent;
    s.twist = sqrt(s.delay_bivector_sq) / fmax(nsq, 1e-300);
    (void)vsq;
    return s;
}

// entropy01 (src/lie_svd_adaptive.rs:375-392) of v[0..n) by the workgroup.
__device__ double block_entropy01(const double* v, int n, double* red) {
    double sum = 0.0;
    for (int i = tix(); i < n; i += kNT) sum += v[i];
    sum = block_sum<kNT>(sum, red);
    if (sum <= 1e-300 || n <= 1) return 1.0;
    double h = 0.0;
    for (int i = tix(); i < n; i += kNT) {
        const double p = v[i] / sum;
        if (p > 0.0) h -= p * log(p);
    }
    h = block_sum<kNT>(h, red);
    return h / fmax(log((double)n), 1e-300);
}

// coeff_var (:348-362) of v[0..n) by the workgroup.
__device__ double block_coeff_var(const double* v, int n, double* red) {
    double sum = 0.0;
    for (int i = tix(); i < n; i += kNT) sum += v[i];
    const double mean = block_sum<kNT>(sum, red) / (double)(n > 1 ? n : 1);
    if (fabs(mean) <= 1e-300) return 0.0;
    double var = 0.0;
    for (int i = tix(); i < n; i += kNT) var += (v[i] - mean) * (v[i] - mean);
    var = block_sum<kNT>(var, red) / (double)(n > 1 ? n : 1);
    return sqrt(var) / fabs(mean);
}

__global__ __launch_bounds__(kNT) void triage_kernel(const TriageItem* items, const double* pool,
                                                     lsvd_adaptive_params_t prm, lsvd_triage_t* out, int max_n) {
    extern __shared__ double sm[];  // row_l1, col_l1, row_l2, col_l2, diag_abs: max_n each
    __shared__ double red[kWaves];
    const TriageItem it = items[blockIdx.x];
    const int n = it.n;
    const double* a = pool + it.off;
    double* row_l1 = sm;
    double* col_l1 = sm + max_n;
    double* row_l2 = sm + 2 * max_n;
    double* col_l2 = sm + 3 * max_n;
    double* diag_abs = sm + 4 * max_n;
    const int tid = tix(), wv = tid / kWave, lane = tid & (kWave - 1);
    // ---- rows, then columns (phase_signature's per-vector health)
    double d_acc = 0.0, g_acc = 0.0, tw_max = 0.0, row_en = 0.0, col_en = 0.0;
    for (int i = wv; i < n; i += kWaves) {
        const VecStats s = wave_vector_phase(a + (int64_t)i * n, n, 1, lane);
        if (lane == 0) {
            row_l1[i] = s.l1;
            row_l2[i] = s.l2;
            diag_abs[i] = fabs(a[(int64_t)i * n + i]);
            d_acc += s.delay_bivector_sq;
            g_acc += s.gradient_sq;
            tw_max = fmax(tw_max, s.twist);
            row_en += s.entropy;
        }
    }
    for (int j = wv; j < n; j += kWaves) {
        const VecStats s = wave_vector_phase(a + j, n, n, lane);
        if (lane == 0) {
            col_l1[j] = s.l1;
            col_l2[j] = s.l2;
            d_acc += s.delay_bivector_sq;
            g_acc += s.gradient_sq;
            tw_max = fmax(tw_max, s.twist);
            col_en += s.entropy;
        }
    }
    d_acc = block_sum<kNT>(d_acc, red);
    g_acc = block_sum<kNT>(g_acc, red);
    tw_max = block_max<kNT>(tw_max, red);
    row_en = block_sum<kNT>(row_en, red);
    col_en = block_sum<kNT>(col_en, red);
    // ---- entries (analyze_matrix's sums, triangular_disbalance, torsion)
    double all_sq = 0.0, diag_sq = 0.0, off_sq = 0.0, sym_sq = 0.0, skew_sq = 0.0, up = 0.0, lo = 0.0;
    const int64_t nn = (int64_t)n * n;
    for (int64_t e = tid; e < nn; e += kNT) {
        const int i = (int)(e / n), j = (int)(e % n);
        const double v = a[e];
        all_sq += v * v;
        if (i == j) diag_sq += v * v;
        else off_sq += v * v;
        if (i < j) {
            const double y = a[(int64_t)j * n + i];
            const double sym = 0.5 * (v + y), skew = 0.5 * (v - y);
            sym_sq += 2.0 * sym * sym;
            skew_sq += 2.0 * skew * skew;
            up += v * v;
            lo += y * y;
        }
    }
    all_sq = block_sum<kNT>(all_sq, red);
    diag_sq = block_sum<kNT>(diag_sq, red);
    off_sq = block_sum<kNT>(off_sq, red);
    sym_sq = block_sum<kNT>(sym_sq, red);
    skew_sq = block_sum<kNT>(skew_sq, red);
    up = block_sum<kNT>(up, red);
    lo = block_sum<kNT>(lo, red);
    // global phase entropy over every entry
    double pent = 0.0;
    if (nn > 1 && all_sq > 1e-300) {
        for (int64_t e = tid; e < nn; e += kNT) {
            const double p = (a[e] * a[e]) / all_sq;
            if (p > 0.0) pent -= p * log(p);
        }
        pent = block_sum<kNT>(pent, red) / log((double)nn);
    }
    // Block4 dual balance over the 4 x 4 diagonal blocks
    double sd2 = 0.0, asd2 = 0.0;
    const double r2 = 0.70710678118654752440;  // 1 / sqrt(2)
    for (int b = tid; b < n / 4; b += kNT) {
        const int st = 4 * b;
        auto skw = [&](int p, int q) {
            return 0.5 * (a[(int64_t)(st + q) * n + st + p] - a[(int64_t)(st + p) * n + st + q]);
        };
        const double w12 = skw(0, 1), w13 = skw(0, 2), w14 = skw(0, 3), w23 = skw(1, 2), w24 = skw(1, 3),
                     w34 = skw(2, 3);
        const double s0 = (w12 + w34) * r2, s1 = (w13 - w24) * r2, s2 = (w14 + w23) * r2;
        const double t0 = (w12 - w34) * r2, t1 = (w13 + w24) * r2, t2 = (w14 - w23) * r2;
        sd2 += s0 * s0 + s1 * s1 + s2 * s2;
        asd2 += t0 * t0 + t1 * t1 + t2 * t2;
    }
    sd2 = block_sum<kNT>(sd2, red);
    asd2 = block_sum<kNT>(asd2, red);
    __syncthreads();  // the norm vectors in LDS are complete
    // ---- the norm vectors
    const double row_cv = block_coeff_var(row_l2, n, red);
    const double col_cv = block_coeff_var(col_l2, n, red);
    double num = 0.0, den = 0.0;
    for (int i = tid; i < n; i += kNT) {
        const double d = row_l1[i] - col_l1[i];
        num += d * d;
        den += row_l1[i] * row_l1[i] + col_l1[i] * col_l1[i];
    }
    num = block_sum<kNT>(num, red);
    den = block_sum<kNT>(den, red);
    const double diag_ent = block_entropy01(diag_abs, n, red);
    const double row_ent = block_entropy01(row_l1, n, red);
    const double col_ent = block_entropy01(col_l1, n, red);
    if (tid != 0) return;
    lsvd_triage_t t{};
    const double norm = fmax(sqrt(all_sq), 1e-300);
    t.n = n;
    t.offdiag_ratio = sqrt(off_sq) / norm;
    t.diagonal_dominance = sqrt(diag_sq) / norm;
    t.row_cv = row_cv;
    t.col_cv = col_cv;
    t.row_col_mismatch = sqrt(num) / fmax(sqrt(den), 1e-300);
    t.symmetry_ratio = sqrt(sym_sq) / fmax(sqrt(sym_sq + skew_sq), 1e-300);
    t.transpose_torsion_ratio = sqrt(skew_sq) / fmax(sqrt(sym_sq + skew_sq), 1e-300);
    t.diagonal_entropy = diag_ent;
    t.row_mass_entropy = row_ent;
    t.col_mass_entropy = col_ent;
    const double cnt = (double)(2 * n > 1 ? 2 * n : 1);
    t.phase_mean_stress = (sqrt(d_acc) + sqrt(g_acc)) / cnt;
    t.phase_max_twist = tw_max;
    t.phase_causal_disbalance = fabs(up - lo) / fmax(up + lo, 1e-300);
    t.phase_entropy_gap = n > 0 ? fabs(row_en / n - col_en / n) : 0.0;
    t.phase_torsion_energy = sqrt(skew_sq);
    {
        const double s = sqrt(sd2), q = sqrt(asd2);
        t.phase_chirality_balance = n / 4 > 0 ? fabs(s - q) / fmax(s + q, 1e-300) : 0.0;
    }
    t.phase_entropy = pent;
    // suspicious_score (src/lie_svd_adaptive.rs:319-339)
    double sc = 0.0;
    sc += fmax(t.offdiag_ratio - 0.55, 0.0) * 0.9;
    sc += fmax(0.35 - t.diagonal_dominance, 0.0) * 0.8;
    sc += fmax(t.row_cv, t.col_cv) * 1.4;
    sc += t.row_col_mismatch * 1.6;
    sc += t.transpose_torsion_ratio * 0.45;
    sc += t.offdiag_ratio * fmax(t.symmetry_ratio - 0.82, 0.0) * 1.8;
    sc += fmax(0.78 - t.diagonal_entropy, 0.0) * 1.2;
    t.suspicious_score = sc;
    t.route = adaptive_route(t, prm);
    out[it.idx] = t;
}

}  // namespace

hipError_t launch_triage(const TriageItem* items, int count, int max_n, const double* pool,
                         const lsvd_adaptive_params_t& p, lsvd_triage_t* out, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    const int mn = max_n > 0 ? max_n : 1;
    const size_t lds = sizeof(double) * 5 * (size_t)mn;
    hipLaunchKernelGGL(triage_kernel, dim3(count), dim3(kNT), lds, st, items, pool, p, out, mn);
    return hipGetLastError();
}

}  // namespace lsvd
