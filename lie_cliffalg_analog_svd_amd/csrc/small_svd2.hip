// small_svd2.hip -- on-chip batched LieSvdSmall for N <= 128, M <= 128.
//
// Same algorithm as small_svd.hip (src/lie_svd_small.rs:296-360), with
// the working set kept on the CU:
//   - matrices are zero-padded to NP = 16 * ceil(N / 16) (the padding stays
//     zero under every step below, and padded Jacobi pairs have a_pq = 0, so
//     the reference's |a_pq| < 1e-300 skip leaves them alone);
//   - Newton-Schulz, P = Q^T R and U_r = Q V run on v_mfma_f64_16x16x4f64 in
//     16-column panels: X lives in LDS, each finished tile of the new iterate
//     / product goes to the staging slot and is copied back after a barrier;
//   - P lives in LDS for the Jacobi sweeps; V lives in registers in the
//     "position space" of the round-robin ring: every layer's pairs sit at
//     fixed register slots (2k, 2k+1) and moving to the next layer is a
//     fixed cyclic shift of the slots (register moves plus one exchange with
//     each neighbouring lane), so V needs neither LDS traffic nor dynamic
//     register indexing.  A sweep is NP - 1 layers, after which the ring is
//     back in its starting arrangement;
//   - Q is parked in a per-workgroup global slot during the sweeps, and the
//     tiles of each product are staged in a second slot (L2 resident), so
//     only one or two MFMA accumulators are live at a time.
// Wide inputs are solved transposed (as in small_svd.hip).
#include <hip/hip_runtime.h>
#include <math.h>

#include "lsvd_device.hpp"
#include "lsvd_kernels.hpp"
#include "lsvd_small_common.hpp"

namespace lsvd {

typedef double d4v __attribute__((ext_vector_type(4)));

// Largest padded size whose Newton-Schulz step may keep Y = 3I - X^T X in
// LDS next to X (the host sizes LDS for it when it fits).
constexpr int kV2YModeMaxNP = 96;
constexpr size_t kV2LdsBudget = 150 * 1024;

// Threads per matrix of the per-size kernels: measured per-matrix latency is
// lowest with 4 (N <= 32), 8 (N <= 64) and 16 (N > 64) waves per matrix.
template <int NP>
constexpr int v2_default_nt() {
    return NP <= 16 ? 64 : (NP <= 32 ? 256 : (NP <= 64 ? 512 : 1024));
}

template <int NP, int NTT = v2_default_nt<NP>()>
struct Cfg2 {
    static constexpr int NT = NTT;
    static constexpr int NW = NT / 64;
    // V in registers: a lane holds VR rows x CH ring positions, a row spread
    // over TPR consecutive lanes.  The lanes of a row group read the
    // layer's (c, s) of their CH / 2 slots once for all VR rows, so the
    // (c, s) traffic per layer is NP^2 / VR x 8 bytes: at NP = 128 two rows
    // per lane (16 lanes a row) halve it for the same 2 CH VR registers
    // (Jacobi cycles -3.4 %; at NP = 96 the same layout measured +7 %).
    static constexpr int TPR = NP <= 16 ? 4 : (NP == 128 ? 16 : 8);  // divides 64
    static constexpr int VR = NP == 128 ? 2 : 1;
    static constexpr int CH = NP / TPR;  // ring positions per lane
    static constexpr int VLANES = NP / VR * TPR;
    static_assert(CH % 2 == 0 && CH >= 4 && NP % VR == 0 && VLANES <= NT, "V ring layout");
    static constexpr int NTILE = NP / 16;
    static constexpr int LDM = NP + 1;
    static constexpr int SLOTS = NP / 2;
    // (c, s) of the layer's slots, grouped by the lanes of a V row (CSG
    // slots each) with an odd group stride in 16-byte units, so the 8
    // groups a wave reads fall in distinct LDS banks
    static constexpr int CSG = CH / 2;
    static constexpr int CSP = (CSG % 2 == 0) ? CSG + 1 : CSG;
    static constexpr int LDP = 17;  // panel leading dimension (16 columns)
    static constexpr int PANEL = (NP * LDP > 16 * LDM) ? NP * LDP : 16 * LDM;
};

template <int NP, int NT = v2_default_nt<NP>()>
struct Sh2 {
    double red[2 * Cfg2<NP, NT>::NW];
    double diag[NP];
    double tau[NP];
    double scr[NP];
    // per layer parity: (c, s oriented to the slot order), slot k at csi(k)
    // (the V rows' reads); 64 more entries take the writes of angle-wave
    // lanes that hold no slot
    double2 cs[2][Cfg2<NP, NT>::TPR * Cfg2<NP, NT>::CSP + 64];
    double2 csp[2][NP / 2 + 64];                   // the same, slot k at k (the P blocks' reads)
    int rank[NP];
    int item;
};

template <int NP>
__device__ __forceinline__ int csi(int k) {
    return (k / Cfg2<NP>::CSG) * Cfg2<NP>::CSP + k % Cfg2<NP>::CSG;
}

__device__ __forceinline__ d4v mfma64(double a, double b, d4v c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// 1/sqrt(x) and 1/x for x within a few binades of 1 (the angle inputs are
// scaled by a power of two first): the hardware estimates (v_rsq_f64,
// v_rcp_f64, relative error ~2^-26) refined by two Newton steps each, a
// shorter dependent chain than the IEEE sqrt and division sequences (whose
// range fix-ups these inputs never need).
__device__ __forceinline__ double rsq_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
    y = y * fma(-h * y, y, 1.5);
    y = y * fma(-h * y, y, 1.5);
    return y;
}

__device__ __forceinline__ double rcp_nr(double x) {
    double y = __builtin_amdgcn_rcp(x);
    y = fma(fma(-x, y, 1.0), y, y);
    y = fma(fma(-x, y, 1.0), y, y);
    return y;
}

// The rotation of one pair (src/lie_svd_small.rs:217-243) in the form of
// jacobi_angle (lsvd_device.hpp): t = sgn |2 a_pq| / (|d| + r), c = (|d| +
// r) / h, s = sgn |2 a_pq| / h with r = sqrt(d^2 + 4 a_pq^2), h = sqrt((|d| +
// r)^2 + 4 a_pq^2), d and 2 a_pq scaled by one power of two so that the
// larger is in [0.5, 1).
__device__ __forceinline__ int jacobi_angle_nr(double x00, double x11, double apq, bool fwd, double& c, double& sv,
                                               double& tv) {
    const bool rot = fabs(apq) >= 1e-300;
    const double app = fwd ? x00 : x11, aqq = fwd ? x11 : x00;
    const double d = aqq - app, a2 = 2.0 * apq;
    const int ex = __builtin_amdgcn_frexp_exp(fmax(fabs(d), fabs(a2)));
    const double ds = ldexp(d, -ex), as = fabs(ldexp(a2, -ex));
    const double x = fma(ds, ds, as * as);
    const double r = x * rsq_nr(x);
    const double den = fabs(ds) + r;
    const double sgn = (d == 0.0 || ((d > 0.0) == (a2 > 0.0))) ? 1.0 : -1.0;  // sgn(theta), theta = +-0 -> +
    const double ih = rsq_nr(fma(den, den, as * as));
    const double tt = sgn * as * rcp_nr(den);
    c = rot ? den * ih : 1.0;
    const double sn = rot ? sgn * as * ih : 0.0;
    sv = fwd ? sn : -sn;
    tv = rot ? (fwd ? tt : -tt) : 0.0;
    return rot ? 1 : 0;
}

// Lane i takes lane i + 1's value (DPP wave_shl:1) / lane i - 1's (wave_shr:1);
// a lane without a source keeps its own.  The whole wave must execute them.
__device__ __forceinline__ double from_next_lane(double x) {
    const int lo = __double2loint(x), hi = __double2hiint(x);
    return __hiloint2double(__builtin_amdgcn_update_dpp(hi, hi, 0x130, 0xf, 0xf, false),
                            __builtin_amdgcn_update_dpp(lo, lo, 0x130, 0xf, 0xf, false));
}

__device__ __forceinline__ double from_prev_lane(double x) {
    const int lo = __double2loint(x), hi = __double2hiint(x);
    return __hiloint2double(__builtin_amdgcn_update_dpp(hi, hi, 0x138, 0xf, 0xf, false),
                            __builtin_amdgcn_update_dpp(lo, lo, 0x138, 0xf, 0xf, false));
}

// One 16 x 16 tile: acc += A (16 x K) B (K x 16); fa(i, k), fb(k, j);
// K is a multiple of 16.
// Operand lanes: A[i][k] at lane (i + 16 k'), B[k][j] at lane (j + 16 k').
// The k each lane group k' reads is ordered for the LDS banks (64 banks of
// 4 bytes; a 64-bit read is served 32 lanes -- k' = 0, 1 or k' = 2, 3 -- per
// pass).  Every operand here is a row (address row * ld + k) or a column
// (address k * ld + col) of a matrix with an odd leading dimension ld, so
// lanes i = 0..15 of one k' cover 32 banks, and the two groups of a pass
// fill the other 32 when their k differ by 16: 2 * 16 * ld = 32 mod 64 for
// the column reads, and 2 * 16 = 32 for the row reads.  With consecutive k
// (k0 + k') the two groups overlapped in 30 of their 32 banks.  So a
// 64-wide block of k is taken as k = q + 16 k' (q < 16), a 32-wide one as
// q + 16 (k' & 1) + 8 (k' >> 1) (q < 8), and a last 16 as k0 + k'.
template <class FA, class FB>
__device__ __forceinline__ d4v tile_mm(int K, const FA& fa, const FB& fb, d4v acc) {
    const int lane = tix() & 63;
    const int i = lane & 15, kk = lane >> 4;
    const int K64 = K & ~63;
    const int o64 = 16 * kk;
#pragma unroll 1
    for (int kb = 0; kb < K64; kb += 64) {
#pragma unroll 8
        for (int q = 0; q < 16; ++q) {
            const int k = kb + q + o64;
            acc = mfma64(fa(i, k), fb(k, i), acc);
        }
        __builtin_amdgcn_iglp_opt(0);  // interleave the LDS reads with the MFMAs
    }
    int k0 = K64;
    if (K - k0 >= 32) {
        const int o32 = k0 + 16 * (kk & 1) + 8 * (kk >> 1);
#pragma unroll 4
        for (int q = 0; q < 8; ++q) acc = mfma64(fa(i, o32 + q), fb(o32 + q, i), acc);
        k0 += 32;
    }
#pragma unroll 4
    for (; k0 < K; k0 += 4) acc = mfma64(fa(i, k0 + kk), fb(k0 + kk, i), acc);
    return acc;
}

// Accumulator element r of a lane sits at row 4 r + lane / 16, column lane % 16
// (measured layout of the FP64 16x16x4 instruction).
__device__ __forceinline__ int trow(int r) { return 4 * r + ((tix() & 63) >> 4); }
__device__ __forceinline__ int tcol() { return tix() & 15; }

// Ring slot -> indices at layer L (circle method over m = NP positions).
__device__ __forceinline__ void ring_pair(int m, int L, int k, int& a, int& b) {
    const int ring = m - 1;
    if (k == 0) {
        a = m - 1;
        b = L % ring;
    } else {
        a = (L + k) % ring;
        b = (L + ring - k) % ring;
    }
}

// Index held by ring position g in the layer-0 arrangement.
__device__ __forceinline__ int ring_idx0(int m, int g) {
    const int k = g >> 1;
    if ((g & 1) == 0) return k == 0 ? m - 1 : k;
    return k == 0 ? 0 : m - 1 - k;
}

// Position a ring position moves to when the ring advances one layer
// (the inverse of the shift applied to V below).
__device__ __forceinline__ int ring_dst(int m, int h) {
    if (h == 0) return 0;
    if (h == 2) return 1;
    if (h == m - 1) return m - 2;
    return (h & 1) ? h + 2 : h - 2;
}

// Ring position of index i in the layer-0 arrangement (inverse of ring_idx0).
__device__ __forceinline__ int ring_pos0(int m, int i) {
    if (i == m - 1) return 0;
    if (i == 0) return 1;
    if (2 * i <= m - 2) return 2 * i;
    return 2 * (m - 1 - i) + 1;
}

struct PhaseClock2 {
    long long* out;
    long long t;
    __device__ PhaseClock2(long long* o) : out(o), t(o && threadIdx.x == 0 ? clock64() : 0) {}
    __device__ __forceinline__ void mark(int ph) {
        if (out && threadIdx.x == 0) {
            const long long now = clock64();
            out[ph] += now - t;
            t = now;
        }
    }
};

template <int NP, int NTT = v2_default_nt<NP>()>
__device__ void solve_item2(const Item& itm, const Pools& P, double* main, double* panel, int panel_rows,
                            double* xs, double* vs, double* ys, Sh2<NP, NTT>& sh, int ns_scaled) {
    using C = Cfg2<NP, NTT>;
    constexpr int NT = C::NT, NW = C::NW, LDM = C::LDM, NTILE = C::NTILE;
    constexpr int TPR = C::TPR, CH = C::CH, SLOTS = C::SLOTS, LDP = C::LDP;
    const int tid = tix(), w = tid >> 6;
    const int m0 = itm.m, n0 = itm.n;
    const bool wide = m0 < n0;
    const int M = wide ? n0 : m0, N = wide ? m0 : n0;
    const double* A = P.a + itm.a_off;
    const View ain{A, n0, wide ? 1 : 0};
    const bool eigh = itm.mode == kModeEigh;
    const double rootn = sqrt((double)N);
    lsvd_info_t info = {};
    const d4v zero4 = {0.0, 0.0, 0.0, 0.0};
    PhaseClock2 clk(P.timing ? P.timing + (size_t)itm.idx * kPhases : nullptr);

    // Tall and wide inputs are not reduced by a QR first: the polar
    // iteration, P = sym(Q^T A) and U = Q V run on the M x N matrix itself.
    // With A = Q_R R this is the reference's iteration on R (:356-359)
    // carried by the orthonormal Q_R, so sigma and V are those of R and U is
    // Q_R U_R, without the two sequential Householder passes.
    const int MP = (M + 15) & ~15, MT = MP >> 4;
    clk.mark(kPhQR);

    int polar_ok = 0, iters = 0;
    if (!eigh) {
        // ---- Newton-Schulz polar (src/lie_svd_small.rs:158-183)
        double sc = 0.0;
        for (int e = tid; e < M * N; e += NT) {
            const double x = ain(e / N, e % N);
            sc += x * x;
        }
        sc = fmax(sqrt(block_sum<NT>(sc, sh.red)), 1e-300);
        for (int e = tid; e < MP * NP; e += NT) {
            const int i = e / NP, j = e % NP;
            main[i * LDM + j] = (i < M && j < N) ? ain(i, j) / sc : 0.0;
        }
        __syncthreads();
        int diverged = 0;
        // Y = alpha I - X^T X fits next to X: compute its upper tiles once per
        // step and update X in place, row tile by row tile (no staging)
        const bool ymode = NP <= kV2YModeMaxNP && panel_rows >= NP;
        NsSchedule sched;
        sched.init(ns_scaled != 0);
        for (int it = 0; it < kPolarMaxIters; ++it) {
            const double alpha = sched.alpha(), cstep = sched.c();
            double dsq = 0.0, gi2 = 0.0;  // ||X_new - X||^2, ||X^T X - I||^2 (N x N part)
            int bad = 0;
            if constexpr (NP <= kV2YModeMaxNP) {
                if (ymode) {
                    constexpr int NTRI = NTILE * (NTILE + 1) / 2;
                    for (int t = w; t < NTRI; t += NW) {
                        int I = 0, rem = t;
                        while (rem >= NTILE - I) {
                            rem -= NTILE - I;
                            ++I;
                        }
                        const int J = I + rem;
                        const d4v g = tile_mm(MP, [&](int i, int k) { return main[k * LDM + I * 16 + i]; },
                                              [&](int k, int j) { return main[k * LDM + J * 16 + j]; }, zero4);
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int row = I * 16 + trow(r), col = J * 16 + tcol();
                            const double y = (row == col ? alpha : 0.0) - g[r];
                            panel[row * LDM + col] = y;
                            panel[col * LDM + row] = y;
                            if (row < N && col < N) {
                                const double e = g[r] - (row == col ? 1.0 : 0.0);
                                gi2 += (I == J ? 1.0 : 2.0) * e * e;
                            }
                        }
                    }
                    __syncthreads();
                    for (int R = w; R < MT; R += NW) {
                        d4v acc[NTILE];
#pragma unroll
                        for (int J = 0; J < NTILE; ++J)
                            acc[J] = tile_mm(NP, [&](int i, int k) { return main[(R * 16 + i) * LDM + k]; },
                                             [&](int k, int j) { return panel[k * LDM + J * 16 + j]; }, zero4);
#pragma unroll
                        for (int J = 0; J < NTILE; ++J)
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                const int o = (R * 16 + trow(r)) * LDM + J * 16 + tcol();
                                const double xn = cstep * acc[J][r];
                                const double d = xn - main[o];
                                main[o] = xn;
                                bad |= !isfinite(xn);
                                dsq += d * d;
                            }
                    }
                }
            }
            if (!ymode) {
                if constexpr (NP >= 80 && NW >= 16) {
                    // X_new tile t = J MT + R (panel J's row tile R) belongs to
                    // wave t % NW, register slot t / NW, until every panel is
                    // done: the products read the old X throughout
                    constexpr int QMAX = (8 * NTILE + NW - 1) / NW;  // MT <= 8
                    d4v xacc[QMAX];
#pragma unroll 1
                    for (int J = 0; J < NTILE; ++J) {
                        // panel = (alpha I - X^T X)[:, J]
                        for (int I = w; I < NTILE; I += NW) {
                            const d4v g = tile_mm(MP, [&](int i, int k) { return main[k * LDM + I * 16 + i]; },
                                                  [&](int k, int j) { return main[k * LDM + J * 16 + j]; }, zero4);
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                const int row = I * 16 + trow(r), col = tcol();
                                panel[row * LDP + col] = (row == J * 16 + col ? alpha : 0.0) - g[r];
                                if (row < N && J * 16 + col < N) {
                                    const double e = g[r] - (row == J * 16 + col ? 1.0 : 0.0);
                                    gi2 += e * e;
                                }
                            }
                        }
                        __syncthreads();
#pragma unroll
                        for (int q = 0; q < QMAX; ++q) {
                            const int t = w + q * NW;
                            if (t >= J * MT && t < (J + 1) * MT) {
                                const int R = t - J * MT;
                                xacc[q] = tile_mm(NP, [&](int i, int k) { return main[(R * 16 + i) * LDM + k]; },
                                                  [&](int k, int j) { return panel[k * LDP + j]; }, zero4);
#pragma unroll
                                for (int r = 0; r < 4; ++r) {
                                    const int row = R * 16 + trow(r), col = J * 16 + tcol();
                                    const double xn = cstep * xacc[q][r];
                                    const double d = xn - main[row * LDM + col];
                                    xacc[q][r] = xn;
                                    bad |= !isfinite(xn);
                                    dsq += d * d;
                                }
                            }
                        }
                        __syncthreads();
                    }
#pragma unroll
                    for (int q = 0; q < QMAX; ++q) {
                        const int t = w + q * NW;
                        if (t < MT * NTILE) {
                            const int J = t / MT, R = t % MT;
#pragma unroll
                            for (int r = 0; r < 4; ++r)
                                main[(R * 16 + trow(r)) * LDM + J * 16 + tcol()] = xacc[q][r];
                        }
                    }
                } else {
#pragma unroll 1
                    for (int J = 0; J < NTILE; ++J) {
                        // panel = (alpha I - X^T X)[:, J]
                        for (int I = w; I < NTILE; I += NW) {
                            const d4v g = tile_mm(MP, [&](int i, int k) { return main[k * LDM + I * 16 + i]; },
                                                  [&](int k, int j) { return main[k * LDM + J * 16 + j]; }, zero4);
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                const int row = I * 16 + trow(r), col = tcol();
                                panel[row * LDP + col] = (row == J * 16 + col ? alpha : 0.0) - g[r];
                                if (row < N && J * 16 + col < N) {
                                    const double e = g[r] - (row == J * 16 + col ? 1.0 : 0.0);
                                    gi2 += e * e;
                                }
                            }
                        }
                        __syncthreads();
                        // X_new[:, J] = c X panel -> staging slot; |X_new - X|^2 on the fly
                        for (int R = w; R < MT; R += NW) {
                            const d4v x = tile_mm(NP, [&](int i, int k) { return main[(R * 16 + i) * LDM + k]; },
                                                  [&](int k, int j) { return panel[k * LDP + j]; }, zero4);
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                const int row = R * 16 + trow(r), col = J * 16 + tcol();
                                const double xn = cstep * x[r];
                                const double d = xn - main[row * LDM + col];
                                ys[row * NP + col] = xn;
                                bad |= !isfinite(xn);
                                dsq += d * d;
                            }
                        }
                        __syncthreads();
                    }
                }
            }
            // a non-finite entry makes the step infinite
            dsq = bad ? INFINITY : dsq;
            block_sum2<NT>(dsq, gi2, sh.red);
            iters = it + 1;
            const double delta = sqrt(dsq);
            if (!(delta <= 1e6 * rootn)) {  // (:169-171): Gram-Schmidt of A
                gram_schmidt_g<NT>(M, N, ain, main, LDM, sh.red);
                diverged = 1;
                break;
            }
            if (!ymode && !(NP >= 80 && NW >= 16)) {
                for (int e = tid; e < MP * NP; e += NT) main[(e / NP) * LDM + e % NP] = ys[e];
                __syncthreads();
            }
            const bool plain = sched.plain();
            sched.after(gi2, delta);
            if (plain && delta < kPolarTol * rootn) break;
        }
        // orthogonality_err (:146-150) of the final X: X^T X - I, N x N part
        {
            double s = 0.0;
            for (int t = w; t < NTILE * NTILE; t += NW) {
                const int I = t / NTILE, J = t % NTILE;
                const d4v g = tile_mm(MP, [&](int i, int k) { return main[k * LDM + I * 16 + i]; },
                                      [&](int k, int j) { return main[k * LDM + J * 16 + j]; }, zero4);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = I * 16 + trow(r), col = J * 16 + tcol();
                    if (row < N && col < N) {
                        const double d = g[r] - (row == col ? 1.0 : 0.0);
                        s += d * d;
                    }
                }
            }
            const double oe = sqrt(block_sum<NT>(s, sh.red));
            // trust test (:174, :178); a diverged iteration is never trusted (:170)
            polar_ok = diverged ? 0 : (oe <= 1e-8 * rootn);
        }
        info.ns_iters = iters;
        clk.mark(kPhNS);

        // ---- P = sym(Q^T A) (:305-309); Q parked in the global slot xs
#pragma unroll 1
        for (int J = 0; J < NTILE; ++J) {
            for (int e = tid; e < MP * 16; e += NT) {
                const int k = e >> 4, j = e & 15;
                panel[k * LDP + j] = (k < M && J * 16 + j < N) ? ain(k, J * 16 + j) : 0.0;
            }
            __syncthreads();
            for (int I = w; I < NTILE; I += NW) {
                const d4v pv = tile_mm(MP, [&](int i, int k) { return main[k * LDM + I * 16 + i]; },
                                       [&](int k, int j) { return panel[k * LDP + j]; }, zero4);
#pragma unroll
                for (int r = 0; r < 4; ++r) ys[(I * 16 + trow(r)) * NP + J * 16 + tcol()] = pv[r];
            }
            __syncthreads();
        }
        for (int e = tid; e < MP * NP; e += NT) {  // each element read, then overwritten, by one thread
            const int o = (e / NP) * LDM + e % NP;
            xs[e] = main[o];
            if (e < NP * NP) main[o] = ys[e];
        }
        __syncthreads();
        for (int e = tid; e < NP * NP; e += NT) {
            const int i = e / NP, j = e % NP;
            if (i < j) {
                const double mm = (main[i * LDM + j] + main[j * LDM + i]) * 0.5;
                main[i * LDM + j] = mm;
                main[j * LDM + i] = mm;
            }
        }
        __syncthreads();
    } else {
        for (int e = tid; e < NP * NP; e += NT) {
            const int i = e / NP, j = e % NP;
            main[i * LDM + j] = (i < N && j < N) ? A[(size_t)i * N + j] : 0.0;
        }
        __syncthreads();
    }
    info.polar_ok = polar_ok;
    clk.mark(kPhP);

    // ---- Jacobi on P (:212-285) in ring-position space.  P is stored
    // permuted so that the pairs of every layer sit at positions (2k, 2k+1):
    // its upper triangle as 2 x 2 blocks (ka <= kb) of the S = NP / 2 slots,
    // packed diagonal by diagonal (block (ka, ka + d) at base(d) + ka, four
    // planes for the entries 00, 01, 10, 11: a layer moves every entry by
    // one block along or across the diagonals, so consecutive lanes of a
    // diagonal read and write consecutive doubles), in two buffers: layer L reads one and writes
    // its rotated blocks straight to the positions layer L + 1 needs them at
    // in the other, so a layer needs one barrier.  Wave 0 (the angle wave)
    // rotates the blocks on and next to the diagonal (kb - ka <= 2) itself:
    // those are all the next layer's diagonal blocks come from, so it forms
    // them in registers and computes layer L + 1's angles while the other
    // waves rotate the rest of P and V for layer L.  V lives in registers
    // in the same position space (TPR lanes per row): a layer is a rotation
    // of register pairs plus a fixed cyclic shift (register moves and one
    // neighbour-lane exchange each way).
    constexpr int S = NP / 2, NBLK = S * (S + 1) / 2;
    constexpr int NR = (S - 3) * (S - 2) / 2;  // blocks with kb - ka >= 3, from base(3) = 3S - 3 on
    constexpr int NRT = NW == 1 ? NT : NT - 64;  // threads taking those
    constexpr int MBR = (NR + NRT - 1) / NRT;
    static_assert(S >= 4, "ring layout");
    auto base = [](int d) { return d * S - (d * (d - 1)) / 2; };  // first block of diagonal d
    // storage offset (within one buffer) of ring position (r, c)
    auto poff = [&](int r, int c) {
        const int lo = r < c ? r : c, hi = r < c ? c : r;
        return (((lo & 1) << 1) | (hi & 1)) * NBLK + base((hi >> 1) - (lo >> 1)) + (lo >> 1);
    };
    // block index -> (diagonal d, ka)
    auto decode = [](int b, int& d, int& ka) {
        d = 0;
        while (b >= S - d) {
            b -= S - d;
            ++d;
        }
        ka = b;
    };
    // 2 x 4 NBLK doubles, then 128 doubles that take the writes of
    // angle-wave lanes holding no block: NP (NP + 2) + 128 <= the LDS area
    double* pb = main;
    constexpr int SINK = 8 * NBLK;
    double ref = 0.0;
    {
        for (int e = tid; e < NP * NP; e += NT) {
            const double x = main[(e / NP) * LDM + e % NP];
            ys[e] = x;
            ref += x * x;
        }
        ref = fmax(sqrt(block_sum<NT>(ref, sh.red)), 1e-300);  // its barriers also order the ys writes
        for (int b = tid; b < NBLK; b += NT) {
            int dd, ka;
            decode(b, dd, ka);
            const int kb = ka + dd;
#pragma unroll
            for (int cmp = 0; cmp < 4; ++cmp) {
                const int g = 2 * ka + (cmp >> 1), hh = 2 * kb + (cmp & 1);
                pb[cmp * NBLK + b] = ys[ring_idx0(NP, g) * NP + ring_idx0(NP, hh)];
            }
        }
        __syncthreads();
    }
    // V rows leave the angle wave alone when the other waves can hold them
    constexpr int VR = C::VR;
    constexpr int VBASE = (NW > 1 && C::VLANES <= NT - 64) ? 64 : 0;
    // rows vrow + rr, rr < VR (vrow >= NP: no V rows)
    const int vrow = tid >= VBASE ? (tid - VBASE) / TPR * VR : NP, vch = (tid - VBASE) % TPR;
    double v[VR][CH];
#pragma unroll
    for (int rr = 0; rr < VR; ++rr)
#pragma unroll
        for (int j = 0; j < CH; ++j) {
            const int idx = ring_idx0(NP, vch * CH + j);
            v[rr][j] = (vrow + rr < N && idx == vrow + rr) ? 1.0 : 0.0;
        }
    // regular blocks (kb - ka >= 3) of this thread: packed (ka, kb, csi(ka),
    // csi(kb)), source block index and the four write offsets
    const int rt = NW == 1 ? tid : tid - 64;
    int rk[MBR], rsrc[MBR], rw01[MBR], rw23[MBR];
#pragma unroll
    for (int i = 0; i < MBR; ++i) {
        const int j = rt + i * NRT;
        rk[i] = -1;
        rsrc[i] = rw01[i] = rw23[i] = 0;
        if (rt >= 0 && j < NR) {
            int dd, ka;
            decode(base(3) + j, dd, ka);
            const int kb = ka + dd;
            const int da0 = ring_dst(NP, 2 * ka), da1 = ring_dst(NP, 2 * ka + 1);
            const int db0 = ring_dst(NP, 2 * kb), db1 = ring_dst(NP, 2 * kb + 1);
            rk[i] = ka | (kb << 8);
            rsrc[i] = base(3) + j;
            rw01[i] = poff(da0, db0) | (poff(da0, db1) << 16);
            rw23[i] = poff(da1, db0) | (poff(da1, db1) << 16);
        }
    }
    // angle wave, lane k < S: diagonal block (k, k), (k, k + 1) for
    // k <= S - 2 and (k - 1, k + 1) for 1 <= k <= S - 2
    const bool awave = w == 0;
    const int k = tid;
    const bool aslot = awave && k < S;
    const bool has1 = aslot && k <= S - 2, has2 = aslot && k >= 1 && k <= S - 2;
    int b0 = 0, b1 = 0, b2 = 0, w0d = 0, w1d = 0, w1a = 0, w1b = 0, w2a = 0, w2b = 0;
    if (aslot) {
        const int p0 = ring_dst(NP, 2 * k), p1 = ring_dst(NP, 2 * k + 1);
        b0 = k;
        w0d = poff(p0, p0) | (poff(p0, p1) << 16);
        w1d = poff(p1, p1);
        if (has1) {
            const int q0 = ring_dst(NP, 2 * k + 2), q1 = ring_dst(NP, 2 * k + 3);
            b1 = S + k;
            w1a = poff(p0, q0) | (poff(p0, q1) << 16);
            w1b = poff(p1, q0) | (poff(p1, q1) << 16);
        }
        if (has2) {
            const int r0 = ring_dst(NP, 2 * k - 2), r1 = ring_dst(NP, 2 * k - 1);
            const int q0 = ring_dst(NP, 2 * k + 2), q1 = ring_dst(NP, 2 * k + 3);
            b2 = 2 * S - 1 + k - 1;
            w2a = poff(r0, q0) | (poff(r0, q1) << 16);
            w2b = poff(r1, q0) | (poff(r1, q1) << 16);
        }
    }
    // slot k's rotation for the current layer and its neighbours' (c, s)
    double ac = 1.0, as = 0.0, at = 0.0, acn = 1.0, asn = 0.0, acp = 1.0, asp = 0.0;
    int arf = 0;
    auto angle = [&](int L, double x00, double x11, double apq) {
        int a, b;
        ring_pair(NP, L, k, a, b);
        arf = jacobi_angle_nr(x00, x11, apq, a < b, ac, as, at);
    };
    auto neighbours = [&]() {  // the whole wave executes the lane moves
        acn = from_next_lane(ac);
        asn = from_next_lane(as);
        acp = from_prev_lane(ac);
        asp = from_prev_lane(as);
    };
    const int max_sweeps = 60 + 20 * N;
    int sweeps = 0, rots = 0, G = 0;
    const int lane = tid & 63;
    // rotate V's own slots with the layer's angles, then shift the ring
    auto v_update = [&](const double2* csl) {
#pragma unroll
        for (int j = 0; j < CH / 2; ++j) {
            const double2 r = csl[vch * C::CSP + j];
#pragma unroll
            for (int rr = 0; rr < VR; ++rr) {
                const double x = v[rr][2 * j], y = v[rr][2 * j + 1];
                v[rr][2 * j] = r.x * x - r.y * y;
                v[rr][2 * j + 1] = r.y * x + r.x * y;
            }
        }
        // in-place ring shift (verified against the circle-method schedule):
        // even slots move down by one pair, odd slots up
        // (DPP lane moves: a row's TPR lanes are consecutive lanes of one
        // wave, and the row ends do not use what crosses them)
#pragma unroll
        for (int rr = 0; rr < VR; ++rr) {
            double* vv = v[rr];
            const double nfe = from_next_lane(vv[0]);       // next lane's first even
            const double plo = from_prev_lane(vv[CH - 1]);  // previous lane's last odd
            const double e2 = vv[2], olast = vv[CH - 1];
#pragma unroll
            for (int j = 0; j + 2 < CH; j += 2) vv[j] = (vch == 0 && j == 0) ? vv[0] : vv[j + 2];
            vv[CH - 2] = (vch == TPR - 1) ? olast : nfe;
#pragma unroll
            for (int j = CH - 1; j >= 3; j -= 2) vv[j] = vv[j - 2];
            vv[1] = (vch == 0) ? e2 : plo;
        }
    };
    if (awave) {  // the first layer's angles
        if (aslot) {
            angle(0, pb[b0], pb[3 * NBLK + b0], pb[NBLK + b0]);
            sh.cs[0][csi<NP>(k)] = make_double2(ac, as);
            sh.csp[0][k] = make_double2(ac, as);
        }
        neighbours();
    }
    __syncthreads();
    for (int sw = 0; sw < max_sweeps && N > 1; ++sw) {
        const double* cur = pb + (G & 1) * 4 * NBLK;
        double off = 0.0;  // upper off-diagonal entries: a diagonal block holds one (plane 01)
        for (int b = tid; b < NBLK; b += NT) {
            const double x01 = cur[NBLK + b];
            off += x01 * x01;
            if (b >= S) {  // not on the diagonal
                const double x00 = cur[b], x10 = cur[2 * NBLK + b], x11 = cur[3 * NBLK + b];
                off += x00 * x00 + x10 * x10 + x11 * x11;
            }
        }
        off = block_sum<NT>(off, sh.red);
        if (!(sqrt(off) >= kJacobiTol * ref)) break;
        ++sweeps;
#pragma unroll 1
        for (int L = 0; L < NP - 1; ++L, ++G) {
            const double* cp = pb + (G & 1) * 4 * NBLK;
            double* np_ = pb + ((G + 1) & 1) * 4 * NBLK;
            const double2* csl = sh.cs[G & 1];
            const double2* cpl = sh.csp[G & 1];
            if (awave) {
                // all reads first (lanes without a block read their diagonal
                // block's slot, which is in range), then the three rotations
                // as independent chains, then the writes of the blocks held
                const int s1 = has1 ? b1 : b0, s2 = has2 ? b2 : b0;
                const double x00 = cp[b0], x01 = cp[NBLK + b0], x11 = cp[3 * NBLK + b0];
                const double c00 = cp[s1], c01 = cp[NBLK + s1], c10 = cp[2 * NBLK + s1], c11 = cp[3 * NBLK + s1];
                const double e00 = cp[s2], e01 = cp[NBLK + s2], e10 = cp[2 * NBLK + s2], e11 = cp[3 * NBLK + s2];
                rots += aslot ? arf : 0;
                // diagonal block (:244-249): the exact 2 x 2 update
                const double n00 = arf ? x00 - at * x01 : x00;
                const double n11 = arf ? x11 + at * x01 : x11;
                const double n01 = arf ? 0.0 : x01;
                // (k, k + 1): rows with slot k, columns with slot k + 1
                const double p00 = ac * c00 - as * c10, p01 = ac * c01 - as * c11;
                const double p10 = as * c00 + ac * c10, p11 = as * c01 + ac * c11;
                const double m00 = acn * p00 - asn * p01, m01 = asn * p00 + acn * p01;
                const double m10 = acn * p10 - asn * p11, d1n11 = asn * p10 + acn * p11;
                // (k - 1, k + 1): rows with slot k - 1, columns with slot k + 1
                const double q00 = acp * e00 - asp * e10, q01 = acp * e01 - asp * e11;
                const double q10 = asp * e00 + acp * e10, q11 = asp * e01 + acp * e11;
                const double r00 = acn * q00 - asn * q01, r01 = asn * q00 + acn * q01;
                const double r10 = acn * q10 - asn * q11, r11 = asn * q10 + acn * q11;
                // one basic block with the angle below, so the scheduler can
                // fill the angle's dependent chain with these stores (offsets
                // from pb; lanes without the block write their sink slot)
                const int nb = ((G + 1) & 1) * 4 * NBLK, sk = SINK + 2 * lane;
                pb[aslot ? nb + (w0d & 0xffff) : sk] = n00;
                pb[aslot ? nb + (w0d >> 16) : sk] = n01;
                pb[aslot ? nb + w1d : sk] = n11;
                pb[has1 ? nb + (w1a & 0xffff) : sk] = m00;
                pb[has1 ? nb + (w1a >> 16) : sk] = m01;
                pb[has1 ? nb + (w1b & 0xffff) : sk] = m10;
                pb[has1 ? nb + (w1b >> 16) : sk] = d1n11;
                pb[has2 ? nb + (w2a & 0xffff) : sk] = r00;
                pb[has2 ? nb + (w2a >> 16) : sk] = r01;
                pb[has2 ? nb + (w2b & 0xffff) : sk] = r10;
                pb[has2 ? nb + (w2b >> 16) : sk] = r11;
                // next (0, 1) <- (0, 2); next (2k, 2k + 1) <- (2k - 1, 2k + 2)
                double apqn = k == 0 ? m00 : r10;
                double e00n = 0.0, e11n = 0.0;
                // the next layer's diagonal block of slot k (positions move
                // as ring_dst: 2k <- 2k + 2, 2k + 1 <- 2k - 1, with the ring's
                // fixed point 0, 1 <- 2 and 2S - 2 <- 2S - 1)
                const double n00_next = from_next_lane(n00);
                const double n11_prev = from_prev_lane(n11);
                const double d1_prev = from_prev_lane(d1n11);
                e00n = k == 0 ? n00 : (k == S - 1 ? n11 : n00_next);
                e11n = k == 0 ? n00_next : n11_prev;
                if (k == S - 1) apqn = d1_prev;  // next (2S-2, 2S-1) <- (2S-3, 2S-1)
                angle(L + 1 < NP - 1 ? L + 1 : 0, e00n, e11n, apqn);  // lanes >= S: unused
                sh.cs[(G + 1) & 1][aslot ? csi<NP>(k) : C::TPR * C::CSP + lane] = make_double2(ac, as);
                sh.csp[(G + 1) & 1][aslot ? k : S + lane] = make_double2(ac, as);
                neighbours();
            }
            // the other blocks: rotate (:244-285), write to the next layer's positions
#pragma unroll
            for (int i = 0; i < MBR; ++i) {
                if (rk[i] < 0) continue;
                const int ca = rk[i] & 255, cb = rk[i] >> 8;
                const int sb = rsrc[i];
                const double b00 = cp[sb], b01 = cp[NBLK + sb], b10 = cp[2 * NBLK + sb], b11 = cp[3 * NBLK + sb];
                const double2 ra = cpl[ca], rb = cpl[cb];
                const double l00 = ra.x * b00 - ra.y * b10, l01 = ra.x * b01 - ra.y * b11;
                const double l10 = ra.y * b00 + ra.x * b10, l11 = ra.y * b01 + ra.x * b11;
                np_[rw01[i] & 0xffff] = rb.x * l00 - rb.y * l01;
                np_[rw01[i] >> 16] = rb.y * l00 + rb.x * l01;
                np_[rw23[i] & 0xffff] = rb.x * l10 - rb.y * l11;
                np_[rw23[i] >> 16] = rb.y * l10 + rb.x * l11;
            }
            if (vrow < NP) v_update(csl);
            __syncthreads();
        }
    }
    clk.mark(kPhJacobi);
    {
        const double* cur = pb + (G & 1) * 4 * NBLK;
        for (int i = tid; i < N; i += NT) {
            const int q = ring_pos0(NP, i);
            sh.diag[i] = cur[((q & 1) ? 3 : 0) * NBLK + (q >> 1)];
        }
    }
    __syncthreads();
    for (int i = tid; i < N; i += NT) {  // rank sort, eigenvalues descending
        const double di = isnan(sh.diag[i]) ? -INFINITY : sh.diag[i];
        int r = 0;
        for (int j = 0; j < N; ++j) {
            const double dj = isnan(sh.diag[j]) ? -INFINITY : sh.diag[j];
            r += (dj > di) || (dj == di && j < i);
        }
        sh.rank[i] = r;
    }
    // V from registers to LDS (index space), zero padding
    for (int e = tid; e < NP * NP; e += NT) main[(e / NP) * LDM + e % NP] = 0.0;
    __syncthreads();
#pragma unroll
    for (int rr = 0; rr < VR; ++rr)
        if (vrow + rr < N) {
#pragma unroll
            for (int j = 0; j < CH; ++j) {
                const int idx = ring_idx0(NP, vch * CH + j);
                if (idx < N) main[(vrow + rr) * LDM + idx] = v[rr][j];
            }
        }
    __syncthreads();
    rots = (int)block_sum<NT>((double)rots, sh.red);
    info.jacobi_sweeps = sweeps;
    info.rotations = rots;

    if (eigh) {
        double* vout = P.u + itm.u_off;
        double* eout = P.s + itm.s_off;
        for (int i = tid; i < N; i += NT) eout[sh.rank[i]] = sh.diag[i];
        for (int e = tid; e < N * N; e += NT) vout[(size_t)(e / N) * N + sh.rank[e % N]] = main[(e / N) * LDM + e % N];
        if (tid == 0 && P.info) P.info[itm.idx] = info;
        __syncthreads();
        return;
    }

    // ---- sigma, Vt (or U for wide), and a global copy of V for the repair
    double* sdst = P.s + itm.s_off;
    for (int i = tid; i < N; i += NT) sdst[sh.rank[i]] = fmax(sh.diag[i], 0.0);
    double* vdst = wide ? P.u + itm.u_off : P.vt + itm.vt_off;
    for (int e = tid; e < N * N; e += NT) {
        const int r = e / N, i = e % N;
        const double x = main[r * LDM + i];
        if (!wide) vdst[(size_t)sh.rank[i] * N + r] = x;
        else vdst[(size_t)r * N + sh.rank[i]] = x;
        vs[(size_t)r * N + sh.rank[i]] = x;  // sorted V
    }

    // ---- U = Q V (:317), columns in sorted order.  Q's rows are staged
    // through the LDS panel, `cr` row tiles at a time; tiles go to ys.
    {
        const int cr_cap = panel_rows / 16 > 0 ? panel_rows / 16 : 1;
#pragma unroll 1
        for (int I0 = 0; I0 < MT; I0 += cr_cap) {
            const int cr = MT - I0 < cr_cap ? MT - I0 : cr_cap;
            for (int e = tid; e < cr * 16 * NP; e += NT) {
                const int i = e / NP, k = e % NP;
                panel[i * LDM + k] = xs[(size_t)(I0 * 16 + i) * NP + k];
            }
            __syncthreads();
            for (int t = w; t < cr * NTILE; t += NW) {
                const int ri = t / NTILE, J = t % NTILE;
                const d4v uv = tile_mm(NP, [&](int i, int k) { return panel[(ri * 16 + i) * LDM + k]; },
                                       [&](int k, int j) { return main[k * LDM + J * 16 + j]; }, zero4);
#pragma unroll
                for (int r = 0; r < 4; ++r) ys[((I0 + ri) * 16 + trow(r)) * NP + J * 16 + tcol()] = uv[r];
            }
            __syncthreads();
        }
    }
    const int rows_cap = MP > NP ? MP : NP;
    for (int e = tid; e < rows_cap * NP; e += NT) main[(e / NP) * LDM + e % NP] = 0.0;
    __syncthreads();
    for (int e = tid; e < M * N; e += NT) {
        const int row = e / N, col = e % N;
        main[row * LDM + sh.rank[col]] = ys[row * NP + col];
    }
    __syncthreads();
    // orthogonality of U, repair if needed (:318-320)
    double oe = 0.0;
    for (int t = w; t < NTILE * NTILE; t += NW) {
        const int I = t / NTILE, J = t % NTILE;
        const d4v g = tile_mm(MP, [&](int i, int k) { return main[k * LDM + I * 16 + i]; },
                              [&](int k, int j) { return main[k * LDM + J * 16 + j]; }, zero4);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = I * 16 + trow(r), col = J * 16 + tcol();
            if (row < N && col < N) {
                const double d = g[r] - (row == col ? 1.0 : 0.0);
                oe += d * d;
            }
        }
    }
    oe = sqrt(block_sum<NT>(oe, sh.red));
    clk.mark(kPhU);
    if (!polar_ok || oe > 1e-8 * rootn) {
        // repair_left_basis_from_right (:185-208): seed = A V S^-1, MGS
        double smax = 0.0;
        for (int i = tid; i < N; i += NT) smax = fmax(smax, fmax(sh.diag[i], 0.0));
        smax = block_max<NT>(smax, sh.red);
        const double stol = 1e-12 * fmax(smax, 1.0);
        for (int i = tid; i < N; i += NT) sh.scr[sh.rank[i]] = fmax(sh.diag[i], 0.0);
        __syncthreads();
        wg_gemm<NT>(M, N, N, ain, View{vs, N, 0}, [&](int i, int j, double acc) {
            const double sj = sh.scr[j];
            xs[(size_t)i * N + j] = sj > stol ? acc / sj : 0.0;
        });
        __syncthreads();
        gram_schmidt_g<NT>(M, N, View{xs, N, 0}, main, LDM, sh.red);
        info.repaired = 1;
    }
    __syncthreads();
    clk.mark(kPhRepair);
    if (!wide) {
        double* u = P.u + itm.u_off;
        for (int e = tid; e < M * N; e += NT) u[(size_t)(e / N) * N + e % N] = main[(e / N) * LDM + e % N];
    } else {
        double* vt = P.vt + itm.vt_off;  // (N x M) = U_t^T
        for (int e = tid; e < M * N; e += NT) vt[(size_t)(e % N) * M + e / N] = main[(e / N) * LDM + e % N];
    }
    clk.mark(kPhOut);
    if (tid == 0 && P.info) {
        info.path = itm.mode == kModeMicroFallback ? 2 : 0;
        info.micro_fallback = itm.mode == kModeMicroFallback ? 1 : 0;
        P.info[itm.idx] = info;
    }
    __syncthreads();
}

template <int NP, int NT = v2_default_nt<NP>()>
__global__ __launch_bounds__(NT) void small_svd2_kernel(const Item* __restrict__ items, const int* count_ptr,
                                                        int* counter, Pools pools, double* gws, int64_t gws_stride,
                                                        int lds_doubles, int ns_scaled) {
    extern __shared__ double dyn_lds[];
    __shared__ Sh2<NP, NT> sh;
    const int count = *count_ptr;
    const int64_t part = gws_stride / 3;
    double* xs = gws + (size_t)blockIdx.x * gws_stride;
    double* vs = xs + part;
    double* ys = vs + part;
    for (;;) {
        if (threadIdx.x == 0) sh.item = atomicAdd(counter, 1);
        __syncthreads();
        const int it = sh.item;
        __syncthreads();
        if (it >= count) break;
        const Item itm = items[it];
        const int M = itm.m > itm.n ? itm.m : itm.n;
        const int MP = (M + 15) & ~15;
        const int rows = MP > NP ? MP : NP;
        double* main = dyn_lds;
        double* panel = dyn_lds + (size_t)rows * Cfg2<NP, NT>::LDM;
        const int panel_rows = (lds_doubles - rows * Cfg2<NP, NT>::LDM) / Cfg2<NP, NT>::LDM;
        solve_item2<NP, NT>(itm, pools, main, panel, panel_rows, xs, vs, ys, sh, ns_scaled);
    }
}

// One persistent launch over the items of every padded size from 64 to 128
// (the queue sorted longest first across sizes): 16 waves per workgroup,
// LDS and global slot sized for the largest item.  Each workgroup takes the
// next item and runs the solver instance of its padded size.
constexpr int kMultiNT = 1024;
union Sh2Multi {
    Sh2<64, kMultiNT> s64;
    Sh2<80, kMultiNT> s80;
    Sh2<96, kMultiNT> s96;
    Sh2<112, kMultiNT> s112;
    Sh2<128, kMultiNT> s128;
};

// The queue holds the items ordered by padded size, largest first (and by
// cost within a size); count_ptr[1 + j] is the end of size class j
// (128, 112, 96, 80, 64).  Workgroups drain the classes in that order, each
// class with its own solver instance, so the largest matrices start first.
// An index fetched past the end of the current class is carried into the
// next class's loop, so every index is solved exactly once.
template <int NP>
__device__ __forceinline__ void multi_drain(const Item* items, int end, int* counter, const Pools& pools,
                                            double* dyn_lds, int lds_doubles, double* xs, double* vs, double* ys,
                                            Sh2<NP, kMultiNT>& sh, int* next_item, int ns_scaled) {
    for (;;) {
        if (*next_item < 0) {  // uniform: next_item is in LDS
            __syncthreads();
            if (threadIdx.x == 0) *next_item = atomicAdd(counter, 1);
            __syncthreads();
        }
        const int it = *next_item;
        if (it >= end) return;  // keep it for a later class
        __syncthreads();
        if (threadIdx.x == 0) *next_item = -1;
        const Item itm = items[it];
        const int M = itm.m > itm.n ? itm.m : itm.n;
        const int MP = (M + 15) & ~15;
        const int rows = MP > NP ? MP : NP;
        double* main = dyn_lds;
        double* panel = dyn_lds + (size_t)rows * Cfg2<NP, kMultiNT>::LDM;
        const int panel_rows = (lds_doubles - rows * Cfg2<NP, kMultiNT>::LDM) / Cfg2<NP, kMultiNT>::LDM;
        solve_item2<NP, kMultiNT>(itm, pools, main, panel, panel_rows, xs, vs, ys, sh, ns_scaled);
        __syncthreads();
    }
}

__global__ __launch_bounds__(kMultiNT) void small_svd2_multi_kernel(const Item* __restrict__ items,
                                                                    const int* count_ptr, int* counter, Pools pools,
                                                                    double* gws, int64_t gws_stride, int lds_doubles,
                                                                    int ns_scaled) {
    extern __shared__ double dyn_lds[];
    __shared__ Sh2Multi sh;
    __shared__ int next_item;
    const int64_t part = gws_stride / 3;
    double* xs = gws + (size_t)blockIdx.x * gws_stride;
    double* vs = xs + part;
    double* ys = vs + part;
    if (threadIdx.x == 0) next_item = -1;
    __syncthreads();
    multi_drain<128>(items, count_ptr[1], counter, pools, dyn_lds, lds_doubles, xs, vs, ys, sh.s128, &next_item, ns_scaled);
    multi_drain<112>(items, count_ptr[2], counter, pools, dyn_lds, lds_doubles, xs, vs, ys, sh.s112, &next_item, ns_scaled);
    multi_drain<96>(items, count_ptr[3], counter, pools, dyn_lds, lds_doubles, xs, vs, ys, sh.s96, &next_item, ns_scaled);
    multi_drain<80>(items, count_ptr[4], counter, pools, dyn_lds, lds_doubles, xs, vs, ys, sh.s80, &next_item, ns_scaled);
    multi_drain<64>(items, count_ptr[5], counter, pools, dyn_lds, lds_doubles, xs, vs, ys, sh.s64, &next_item, ns_scaled);
}

hipError_t launch_small2_multi(const SmallLaunch& L, hipStream_t st) {
    if (L.lds_bytes > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&small_svd2_multi_kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)L.lds_bytes);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(small_svd2_multi_kernel, dim3(L.grid), dim3(kMultiNT), L.lds_bytes, st, L.items, L.count_ptr,
                       L.counter, L.pools, L.gws, L.gws_stride, (int)(L.lds_bytes / sizeof(double)), L.ns_scaled);
    return hipGetLastError();
}

int small2_multi_blocks_per_cu(size_t lds) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, small_svd2_multi_kernel, kMultiNT, lds) != hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

template <int NP, int NT = v2_default_nt<NP>()>
static hipError_t launch2(const SmallLaunch& L, hipStream_t st) {
    if (L.lds_bytes > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&small_svd2_kernel<NP, NT>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)L.lds_bytes);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((small_svd2_kernel<NP, NT>), dim3(L.grid), dim3(NT), L.lds_bytes, st, L.items, L.count_ptr,
                       L.counter, L.pools, L.gws, L.gws_stride, (int)(L.lds_bytes / sizeof(double)), L.ns_scaled);
    return hipGetLastError();
}

template <int NP, int NT = v2_default_nt<NP>()>
static int blocks2(size_t lds) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, small_svd2_kernel<NP, NT>, NT, lds) != hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

#define LSVD_V2_SIZES(X) X(16) X(32) X(48) X(64) X(80) X(96) X(112) X(128)
// Sizes with a second instance at twice the default threads (the autotuner's
// choice per bucket).
#define LSVD_V2_WIDE_SIZES(X) X(16) X(32) X(48) X(64)

int v2_threads(int np) {
    switch (np) {
#define X(np_) \
    case np_:  \
        return Cfg2<np_>::NT;
        LSVD_V2_SIZES(X)
#undef X
    }
    return 0;
}

bool v2_has_wide(int np) {
    switch (np) {
#define X(np_) \
    case np_:  \
        return true;
        LSVD_V2_WIDE_SIZES(X)
#undef X
    }
    return false;
}

// LDS of one workgroup: the matrix area (max(16 ceil(M/16), NP) rows of
// LDM) and the panel (the larger of a K x 16 operand panel and 16 rows of
// LDM).
size_t v2_lds_bytes(int np, int M) {
    const int mp = (M + 15) & ~15;
    const int rows = mp > np ? mp : np;
    const int ldm = np + 1;
    const int panel = rows * 17 > 16 * ldm ? rows * 17 : 16 * ldm;
    const int ypanel = panel > np * ldm ? panel : np * ldm;  // room for Y (NP x LDM)
    const size_t with_y = sizeof(double) * ((size_t)rows * ldm + ypanel);
    if (np <= kV2YModeMaxNP && with_y <= kV2LdsBudget) return with_y;
    return sizeof(double) * ((size_t)rows * ldm + panel);
}

// Global slot of one workgroup: three parts (parked Q, sorted V, staging)
// of max(16 ceil(M/16), NP) x NP doubles each, 256-byte aligned.
int64_t v2_slot_doubles(int np, int M) {
    const int mp = (M + 15) & ~15;
    const int rows = mp > np ? mp : np;
    const int64_t part = ((int64_t)rows * np + 31) / 32 * 32;
    return 3 * part;
}

hipError_t launch_small2(int np, bool wide, const SmallLaunch& L, hipStream_t st) {
    if (wide) {
        switch (np) {
#define X(np_) \
    case np_:  \
        return launch2<np_, 2 * v2_default_nt<np_>()>(L, st);
            LSVD_V2_WIDE_SIZES(X)
#undef X
        }
        return hipErrorInvalidValue;
    }
    switch (np) {
#define X(np_) \
    case np_:  \
        return launch2<np_>(L, st);
        LSVD_V2_SIZES(X)
#undef X
    }
    return hipErrorInvalidValue;
}

int small2_blocks_per_cu(int np, bool wide, size_t lds) {
    if (wide) {
        switch (np) {
#define X(np_) \
    case np_:  \
        return blocks2<np_, 2 * v2_default_nt<np_>()>(lds);
            LSVD_V2_WIDE_SIZES(X)
#undef X
        }
        return 1;
    }
    switch (np) {
#define X(np_) \
    case np_:  \
        return blocks2<np_>(lds);
        LSVD_V2_SIZES(X)
#undef X
    }
    return 1;
}

}  // namespace lsvd
