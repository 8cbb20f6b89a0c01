// onesided.hip -- batched thin SVD of M x N matrices (N <= 128, M <= 128
// after transposing wide inputs) by Householder QR and one-sided (Hestenes)
// Jacobi on R^T, with the whole Jacobi state in registers.
//
// The same factorisation as LieSvdSmall::solve_rectangular
// (src/lie_svd_small.rs:340-360): A = Q R, then the SVD of the small factor,
// U = Q U_r.  The small factor's SVD is computed by one-sided Jacobi on
// X = R^T instead of Newton-Schulz + eigh_jacobi_full (:296-323): with
// X W = B (columns pairwise orthogonal to working precision), sigma_j =
// ||b_j||, V_A = B diag(sigma)^-1 and U_A = Q W.  R^T is the classical
// preconditioned form of one-sided Jacobi (Drmac and Veselic, "New fast and
// accurate Jacobi SVD algorithm", SIAM J. Matrix Anal. Appl. 29, 2008): its
// columns are graded, so the sweeps converge in about as many sweeps as the
// reference's eigh_jacobi_full while no polar iteration runs at all.
//
// Rotation rule: the pair (p, q) with Gram entries a = ||x_p||^2, b =
// ||x_q||^2, g = x_p . x_q is rotated when g^2 > tol^2 a b, tol = sqrt(N)
// 2^-52 (the relative test of LAPACK's dgesvj), by the Jacobi rotation of
// the 2 x 2 Gram matrix (the reference's Rutishauser form, shared with the
// two-sided kernels).  The iteration ends when a whole sweep's worth of
// consecutive layers applies no rotation, or at the reference's sweep cap
// 60 + 20 N.
//
// Parallel ordering: odd-even ("caterpillar") layers over NE positions (N
// rounded up to whole lanes' positions) -- even layers pair (2k, 2k+1), odd
// layers (2k+1, 2k+2) -- and the two columns of every pair always trade
// positions (rotated or not), so the movement of the columns is independent
// of the data: any NE consecutive layers meet every pair exactly once, and
// the index held by a position
// after G layers has a closed form (index_at).  In registers a trade is
// free (the rotation writes its two outputs to swapped registers); only the
// pair that straddles two lanes in an odd layer needs one neighbour-lane
// move of each column.
//
// Register layout (per workgroup of NW waves, one matrix): positions are
// split over TPR consecutive lanes (CH = NP / TPR per lane), rows over the
// lane groups ("rowgroups"), XR rows of X and the same XR rows of W per
// lane.  Per layer a lane forms the partial Gram entries of its P = CH / 2
// pairs over its rows; the rowgroups of a wave are reduced with the gfx950
// permlane swaps (reduce-scatter over lane bits 5 and 4, so each lane ends
// up owning the sums of PP / 4 pairs, PP = P rounded up to a power of two)
// and DPP (bit 3), the waves through LDS; each lane computes the angles of
// the pairs it owns, and the swaps gather the (c, s) back.
#include <hip/hip_runtime.h>
#include <math.h>

#include <type_traits>

#include "lsvd_device.hpp"
#include "lsvd_kernels.hpp"
#include "lsvd_small_common.hpp"

namespace lsvd {

// ---- configuration -------------------------------------------------------

// Launch shape per padded size NP: TPR lanes per row, NW waves, CH
// positions per lane.  The layout holds NL = TPR * CH >= NP positions; with
// NL > NP the positions from NP on are zero columns that never meet the
// real ones (the odd-layer pair across position NP - 1 is dead).
#define LSVD_OS_CFG_16 8, 1, 2
#ifndef LSVD_OS_CFG_32
#define LSVD_OS_CFG_32 8, 2, 4
#endif
#ifndef LSVD_OS_CFG_48
#define LSVD_OS_CFG_48 32, 4, 2
#endif
#ifndef LSVD_OS_CFG_64
#define LSVD_OS_CFG_64 32, 4, 2
#endif
#ifndef LSVD_OS_CFG_80
#define LSVD_OS_CFG_80 64, 8, 2
#endif
#ifndef LSVD_OS_CFG_96
#define LSVD_OS_CFG_96 64, 8, 2
#endif
#ifndef LSVD_OS_CFG_112
#define LSVD_OS_CFG_112 64, 8, 2
#endif
#ifndef LSVD_OS_CFG_128
#define LSVD_OS_CFG_128 64, 8, 2
#endif
template <int TPR_, int NW_, int CH_>
struct OsShape {
    static constexpr int TPR = TPR_, NW = NW_, CH = CH_;
};
template <int NP>
struct OsDefault;
template <>
struct OsDefault<16> : OsShape<LSVD_OS_CFG_16> {};
template <>
struct OsDefault<32> : OsShape<LSVD_OS_CFG_32> {};
template <>
struct OsDefault<48> : OsShape<LSVD_OS_CFG_48> {};
template <>
struct OsDefault<64> : OsShape<LSVD_OS_CFG_64> {};
template <>
struct OsDefault<80> : OsShape<LSVD_OS_CFG_80> {};
template <>
struct OsDefault<96> : OsShape<LSVD_OS_CFG_96> {};
template <>
struct OsDefault<112> : OsShape<LSVD_OS_CFG_112> {};
template <>
struct OsDefault<128> : OsShape<LSVD_OS_CFG_128> {};

constexpr int ilog2c(int x) { return x <= 1 ? 0 : 1 + ilog2c(x / 2); }
constexpr int pow2up(int x) { return x <= 1 ? 1 : 2 * pow2up((x + 1) / 2); }

template <int NP, int TPR, int NW, int CHH>
struct OsCfg {
    static constexpr int NT = 64 * NW;
    static constexpr int CH = CHH;             // positions per lane
    static constexpr int NL = TPR * CH;        // positions in the layout (>= NP)
    static constexpr int P = CH / 2;           // pairs per lane and layer
    static constexpr int PP = pow2up(P);       // padded to a power of two for the reduction
    static constexpr int RGW = 64 / TPR;       // rowgroups per wave
    static constexpr int LB = ilog2c(RGW);     // lane bits of the rowgroup (the top LB of 6)
    static constexpr int RG = NW * RGW;
    static constexpr int XR = (NP + RG - 1) / RG;  // X rows (and W rows) per lane
    static constexpr int SCMAX = LB < 2 ? LB : 2;  // lane bits 5 and 4, where they are rowgroup bits
    static constexpr int SC = ilog2c(PP) < SCMAX ? ilog2c(PP) : SCMAX;  // scatter steps (bits 5, 4)
    static constexpr int NA = PP >> SC;        // pairs a lane owns after the scatter
    static constexpr int S = NL / 2;           // pairs in the layout
    static constexpr int LDA = NP + 1;
    static_assert(CH % 2 == 0 && CH >= 2 && NL >= NP && NP % CH == 0, "positions per lane");
    static_assert(TPR == 8 || TPR == 16 || TPR == 32 || TPR == 64, "rowgroups of whole lane bit groups");
    static_assert(XR >= 1 && 3 * NW * NL <= NP * LDA, "rows per lane; the sweeps' partial sums fit the matrix area");
};

// Column norms (of the input scaled to ||A||_F = 1) at or below this are
// treated as zero: the Gram products of two such columns underflow, so the
// rotation test cannot orthogonalise them.
constexpr double kOsTiny = 1e-140;

struct OsShared {
    double red[2 * 16];
    int rank[128];
    double sig[128];     // normalised singular values by index
    double tw[128];      // tau scal^2 per column (the QR's update coefficient)
    double scal[128];    // 1 / (alpha - beta) per column (reflector tails = scal x)
    double tau[128];
    double tau_a[128];   // A's tau while Y is factored (the W-free sizes)
    int item;
    int flag;
};

// ---- lane exchanges ------------------------------------------------------

// a' = [a lanes 0-31 | b lanes 0-31], b' = [a lanes 32-63 | b lanes 32-63]
__device__ __forceinline__ void xch32(double& a, double& b) {
    const int al = __double2loint(a), ah = __double2hiint(a);
    const int bl = __double2loint(b), bh = __double2hiint(b);
    const auto l = __builtin_amdgcn_permlane32_swap(al, bl, false, false);
    const auto h = __builtin_amdgcn_permlane32_swap(ah, bh, false, false);
    a = __hiloint2double(h[0], l[0]);
    b = __hiloint2double(h[1], l[1]);
}

// rows of 16 lanes: a' = [a r0, b r0, a r2, b r2], b' = [a r1, b r1, a r3, b r3]
__device__ __forceinline__ void xch16(double& a, double& b) {
    const int al = __double2loint(a), ah = __double2hiint(a);
    const int bl = __double2loint(b), bh = __double2hiint(b);
    const auto l = __builtin_amdgcn_permlane16_swap(al, bl, false, false);
    const auto h = __builtin_amdgcn_permlane16_swap(ah, bh, false, false);
    a = __hiloint2double(h[0], l[0]);
    b = __hiloint2double(h[1], l[1]);
}

// lane l ^ 8's value (DPP row_ror:8 within rows of 16)
__device__ __forceinline__ double from_xor8(double x) {
    const int lo = __double2loint(x), hi = __double2hiint(x);
    return __hiloint2double(__builtin_amdgcn_mov_dpp(hi, 0x128, 0xf, 0xf, true),
                            __builtin_amdgcn_mov_dpp(lo, 0x128, 0xf, 0xf, true));
}

// lane l + 1's value (wave_shl:1), lane l - 1's (wave_shr:1); a lane
// without a source (63, resp. 0) reads 0.  No `old` operand, so the
// compiler needs no copy of the destination first.
__device__ __forceinline__ double os_next0(double x) {
    return __hiloint2double(__builtin_amdgcn_mov_dpp(__double2hiint(x), 0x130, 0xf, 0xf, true),
                            __builtin_amdgcn_mov_dpp(__double2loint(x), 0x130, 0xf, 0xf, true));
}
__device__ __forceinline__ double os_prev0(double x) {
    return __hiloint2double(__builtin_amdgcn_mov_dpp(__double2hiint(x), 0x138, 0xf, 0xf, true),
                            __builtin_amdgcn_mov_dpp(__double2loint(x), 0x138, 0xf, 0xf, true));
}

// lane l + 1's value (wave_shl:1); lane 63 gets `old`
__device__ __forceinline__ double os_from_next(double x, double old) {
    return __hiloint2double(
        __builtin_amdgcn_update_dpp(__double2hiint(old), __double2hiint(x), 0x130, 0xf, 0xf, false),
        __builtin_amdgcn_update_dpp(__double2loint(old), __double2loint(x), 0x130, 0xf, 0xf, false));
}

// lane l - 1's value (wave_shr:1); lane 0 gets `old`
__device__ __forceinline__ double os_from_prev(double x, double old) {
    return __hiloint2double(
        __builtin_amdgcn_update_dpp(__double2hiint(old), __double2hiint(x), 0x138, 0xf, 0xf, false),
        __builtin_amdgcn_update_dpp(__double2loint(old), __double2loint(x), 0x138, 0xf, 0xf, false));
}

// Sum over the lanes that differ in bit 5 / bit 4 / bit 3 (every lane gets
// the same bits: a + b is evaluated as lower-lane value + upper-lane value,
// or commutatively).
__device__ __forceinline__ double bfly32(double v) {
    double a = v, b = v;
    xch32(a, b);
    return a + b;
}
__device__ __forceinline__ double bfly16(double v) {
    double a = v, b = v;
    xch16(a, b);
    return a + b;
}
__device__ __forceinline__ double bfly8(double v) { return v + from_xor8(v); }

// Index held by position g after G layers of the odd-even ordering over NP
// positions (position p holds index p before the first layer).  Unfolded,
// every column walks a cycle of 2 NP virtual positions one step per layer
// (virtual v is position v going up, position 2 NP - 1 - v going down;
// even positions start going up); the column at g is the one whose virtual
// position has the parity of G.
__device__ __forceinline__ int os_index_at(int NP, int G, int g) {
    const int two = 2 * NP;
    const int v = ((g & 1) == (G & 1)) ? g : two - 1 - g;
    const int v0 = ((v - G) % two + two) % two;
    return v0 < NP ? v0 : two - 1 - v0;
}

// 1/sqrt(x) for x within a few binades of 1: the hardware estimate y
// (v_rsq_f64, relative error ~2^-26) refined by one third-order step,
// 1/sqrt(x) = y (1 - e)^(-1/2) = y (1 + e/2 + 3 e^2/8 + O(e^3)) with
// e = 1 - x y^2 (four dependent operations instead of two Newton steps' six).
__device__ __forceinline__ double os_rsq(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-x * y, y, 1.0);
    return fma(y * e, fma(e, 0.375, 0.5), y);
}

// The Jacobi rotation of the 2 x 2 Gram matrix [[a, g], [g, b]]
// (src/lie_svd_small.rs:226-243 in the form of jacobi_angle): columns
// (x_p, x_q) -> (c x_p - s x_q, s x_p + c x_q) makes them orthogonal.  Its
// inputs are scaled by a power of two so that max(|b - a|, |2 g|) is in
// [0.5, 1); r and 1 / h from reciprocal square roots.  With WANT_T it also
// returns t = s / c (from the hardware reciprocal refined by one third-order
// step): the rotated Gram diagonal is then a - t g (the column c x_p - s x_q)
// and b + t g (s x_p + c x_q).
template <bool WANT_T = false>
__device__ __forceinline__ void os_angle(double a, double b, double g, double& c, double& s, double* t = nullptr) {
    const double d = b - a, a2 = 2.0 * g;
    const int ex = __builtin_amdgcn_frexp_exp(fmax(fabs(d), fabs(a2)));
    const double ds = ldexp(d, -ex), ga = ldexp(a2, -ex);
    // s carries sign(d) sign(g) (+ for d = +0): g's scaled value with d's
    // sign bit folded into its own
    const double sas = __hiloint2double(__double2hiint(ga) ^ (__double2hiint(ds) & (int)0x80000000),
                                        __double2loint(ga));
    const double x = fma(ds, ds, ga * ga);  // in [0.25, 2)
    const double r = x * os_rsq(x);
    const double den = fabs(ds) + r;  // in [0.5, 2.5)
    const double ih = os_rsq(fma(den, den, ga * ga));  // in [0.25, 6)
    c = den * ih;
    s = sas * ih;
    if constexpr (WANT_T) {
        // 1/den = y / (1 - e) = y (1 + e + e^2 + O(e^3)), e = 1 - den y
        const double y = __builtin_amdgcn_rcp(den);
        const double e = fma(-den, y, 1.0);
        *t = sas * fma(y, fma(e, e, e), y);
    }
}

// Incremental column norms: a rotated column whose squared norm falls below
// this fraction of its old value has lost that many bits to cancellation,
// so the next two layers form the Gram diagonal from the columns again.
constexpr double kOsRefresh = 0.0625;

// ---- Householder QR and the application of Q (registers) ------------------

// Column-resident layout of an up to 128-row matrix over the workgroup:
// thread t holds column t / 4 (of NT / 4 >= NP column slots), rows
// t % 4 + 4 i (i < 32), so the four threads of a column are one quad of a
// wave and its sums are quad reductions (DPP), never LDS round trips.
constexpr int kOsQK = 4;               // threads per column
constexpr int kOsQR = 128 / kOsQK;     // row slots per thread
constexpr int kOsRows = 128;           // rows of the LDS matrix area

// From padded size 80 on the sweeps rotate X only (no W): V = B / sigma as
// before, and U comes from Y = R V diag(sigma)^-1 (= W up to rounding, with
// column errors eps sigma_max / sigma_i), orthonormalised by a second
// register QR in sigma order, U = Q_A [Q_Y D; 0], D the signs of R_Y's
// diagonal.  The matrix then lives in registers and the global slot, and
// the LDS area is only the sweeps' partial sums (or a staged chunk), so two
// matrices of 8 waves share a CU (128 registers a lane).
constexpr int kOsWFreeMinNP = 80;
constexpr int kOsKC = 8;    // k per chunk of the product R V
constexpr int kOsCKF = 8;   // reflectors per staged chunk when W-free
constexpr int os_area_doubles(int np) { return np >= kOsWFreeMinNP ? 3 * 8 * 128 : kOsRows * (np + 1); }
constexpr int os_min_waves(int np) { return np >= kOsWFreeMinNP ? 4 : 1; }
// Y's QR and Q_Y keep the pivot / reflector values of a step live in
// registers up to 24 row slots (NP <= 96); above that they read them again
template <int NP>
constexpr bool kOsYTight = NP / kOsQK > 24;

template <int CTRL>
__device__ __forceinline__ double os_dpp(double x) {
    return __hiloint2double(__builtin_amdgcn_mov_dpp(__double2hiint(x), CTRL, 0xf, 0xf, true),
                            __builtin_amdgcn_mov_dpp(__double2loint(x), CTRL, 0xf, 0xf, true));
}

// Sum over the four lanes of a quad (quad_perm [1,0,3,2], then [2,3,0,1]);
// every lane gets the same bits.
__device__ __forceinline__ double os_quad_sum(double v) {
    v = v + os_dpp<0xB1>(v);
    return v + os_dpp<0x4E>(v);
}

// A workgroup barrier that orders LDS only: global stores still in flight
// (the W-free QR writes R and the reflector tails to the global slot, read
// only after the factorisation) are not waited for.
__device__ __forceinline__ void os_lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// p[0] + p[st] + ... + p[(n - 1) st] as a pairwise tree
template <int n>
__device__ __forceinline__ double os_tree_sum(const double* p, int st) {
    if constexpr (n == 1) {
        return p[0];
    } else {
        constexpr int h = n / 2;
        return os_tree_sum<h>(p, st) + os_tree_sum<n - h>(p + (size_t)h * st, st);
    }
}

// Sum over the wave, valid in lane 0 (the other lanes hold partial sums):
// quads by DPP quad_perm, rows of 16 by row_ror 4 and 8, then the swaps
// across bits 4 and 5.
__device__ __forceinline__ double os_wave_sum0(double v) {
    v = os_quad_sum(v);
    v = v + os_dpp<0x124>(v);
    v = v + os_dpp<0x128>(v);
    v = bfly16(v);
    return bfly32(v);
}

// Householder QR of the M x N matrix held in registers (a[i] = row
// t % 4 + 4 i of column t / 4), in the LAPACK dlarfg convention: H_j = I -
// tau_j v_j v_j^T, v_j = 1 at row j.  On return `am` (leading dimension
// lda) holds R on and above the diagonal and the reflector tails v_j below
// it, sh.tau the tau_j.
//
// Step j needs the reflector of column j in every thread.  The quad of
// column j writes the column into vb[j % 2] (128 rows), reads alpha = a_jj
// back, zeroes the rows <= j (the whole wave, after the quad's stores, in
// one wave's program order), forms ||x||^2 of the rows > j from the buffer
// (two rows a lane, a wave sum), then beta, tau and 1 / (alpha - beta) =
// scal, and stores alpha - beta at row j.  With that buffer w (w_r = 0 for r < j, alpha - beta at j, x_r
// below), y_c = sum_r w_r a_rc = (alpha - beta) v_j^T a_c, and column c's
// update is a_c -= tau scal^2 y_c w: one pass over the column's registers
// for the sum, one for the update, no masks.  The quad of column j + 1 builds
// the next buffer right after its own update, so a column costs one barrier.
template <int NT, bool TIGHT = false, int QS = kOsQR>
__device__ __forceinline__ void os_qr_regs(int M, int N, double (&a)[QS], double* am, int lda, double* vb, OsShared& sh) {
    const int tid = tix(), c = tid >> 2, ch = tid & 3, lane = tid & 63;
    // the reflector buffer of column p, built by the wave that holds it
    auto pivot = [&](int p) {
        if ((tid >> 6) != (p >> 4)) return;  // wave-uniform
        double* w = vb + (p & 1) * kOsRows;
        if (c == p)
#pragma unroll
            for (int i = 0; i < QS; ++i) w[ch + kOsQK * i] = a[i];
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        const double alpha = w[p];
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        if (lane <= p) w[lane] = 0.0;
        if (lane + 64 <= p) w[lane + 64] = 0.0;
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        // ||x||^2 of the rows > p over the whole wave (two rows a lane)
        const double x0 = w[lane], x1 = w[lane + 64];
        const double xnorm2 = os_wave_sum0(fma(x0, x0, x1 * x1));
        if (lane == 0) {
            double tau = 0.0, beta = alpha, scal = 0.0;
            if (xnorm2 > 0.0) {
                beta = -copysign(sqrt(alpha * alpha + xnorm2), alpha);
                tau = (beta - alpha) / beta;
                scal = 1.0 / (alpha - beta);
            }
            w[p] = alpha - beta;
            sh.tau[p] = tau;
            sh.tw[p] = tau * scal * scal;
            sh.scal[p] = scal;
            am[(size_t)p * lda + p] = beta;
        }
    };
    if (N > 0) pivot(0);
    for (int j = 0; j < N; ++j) {
        os_lds_barrier();
        const double* w = vb + (j & 1) * kOsRows;
        // the tail of reflector j into am's lower part (v = scal x)
        for (int r = j + 1 + tid; r < M; r += NT) am[(size_t)r * lda + j] = w[r] * sh.scal[j];
        if (c > j && c < N) {  // quad-uniform
            double y0 = 0.0, y1 = 0.0, y2 = 0.0, y3 = 0.0;
#pragma unroll
            for (int i = 0; i < QS; i += 4) {
                y0 = fma(w[ch + kOsQK * i], a[i], y0);
                y1 = fma(w[ch + kOsQK * (i + 1)], a[i + 1], y1);
                y2 = fma(w[ch + kOsQK * (i + 2)], a[i + 2], y2);
                y3 = fma(w[ch + kOsQK * (i + 3)], a[i + 3], y3);
            }
            const double k = sh.tw[j] * os_quad_sum((y0 + y1) + (y2 + y3));
            // a 128-register budget: read w again rather than keep 32 more
            // values live across the sum
            if constexpr (TIGHT) __asm__ __volatile__("" ::: "memory");
#pragma unroll
            for (int i = 0; i < QS; ++i) a[i] = fma(-k, w[ch + kOsQK * i], a[i]);
        }
        if (j + 1 < N) pivot(j + 1);
    }
    __syncthreads();
    // R above the diagonal (the diagonal is in am already)
#pragma unroll
    for (int i = 0; i < QS; ++i) {
        const int r = ch + kOsQK * i;
        if (r < c && c < N) am[(size_t)r * lda + c] = a[i];
    }
}

// u <- H_0 H_1 ... H_{N-1} u for the M-row columns in u (the QR's register
// layout): the reflector tails are read from refl (rows > j of column j,
// leading dimension ldr), tau from `tau`.  The columns are independent, so
// the reflectors are staged in LDS CK at a time (`stg`: 128 x (CK + 1)
// doubles, zero above row j, 1 at row j) and each column's quad applies a
// chunk without a barrier.  TIGHT (a 128-register budget) reads the staged
// values again for the update instead of keeping 32 of them live.
template <int NT, int CK, bool TIGHT, int QS = kOsQR>
__device__ __forceinline__ void os_apply_refl(int M, int N, double (&u)[QS], const double* refl, int ldr, const double* tau,
                              double* stg) {
    constexpr int SL = CK + 1;
    const int tid = tix(), c = tid >> 2, ch = tid & 3;
    for (int j0 = ((N - 1) / CK) * CK; j0 >= 0; j0 -= CK) {
        __syncthreads();  // stg free (the previous chunk applied, or its earlier use done)
        for (int e = tid; e < kOsQK * QS * CK; e += NT) {
            const int r = e / CK, jj = e % CK, j = j0 + jj;
            stg[r * SL + jj] = (r > j && r < M && j < N) ? refl[(size_t)r * ldr + j] : (r == j ? 1.0 : 0.0);
        }
        __syncthreads();
        if (c < N) {  // quad-uniform
            const int jn = N - j0 < CK ? N - j0 : CK;
            for (int jj = jn - 1; jj >= 0; --jj) {
                const double t = tau[j0 + jj];
                if (t == 0.0) continue;  // uniform
                const double* vj = stg + ch * SL + jj;  // row ch + 4 i: vj[4 i SL]
                double y0 = 0.0, y1 = 0.0, y2 = 0.0, y3 = 0.0;
#pragma unroll
                for (int i = 0; i < QS; i += 4) {
                    y0 = fma(vj[kOsQK * i * SL], u[i], y0);
                    y1 = fma(vj[kOsQK * (i + 1) * SL], u[i + 1], y1);
                    y2 = fma(vj[kOsQK * (i + 2) * SL], u[i + 2], y2);
                    y3 = fma(vj[kOsQK * (i + 3) * SL], u[i + 3], y3);
                }
                const double k = t * os_quad_sum((y0 + y1) + (y2 + y3));
                if constexpr (TIGHT) __asm__ __volatile__("" ::: "memory");
#pragma unroll
                for (int i = 0; i < QS; ++i) u[i] = fma(-k, vj[kOsQK * i * SL], u[i]);
            }
        }
    }
}

// U (M x N) from the column-resident registers to row-major output, or its
// transpose (N x M) for a wide input.
template <int QS>
__device__ __forceinline__ void os_write_u(const double (&u)[QS], int M, int N, double* ud, bool wide) {
    const int qc = tix() >> 2, qch = tix() & 3;
    if (qc < N)
#pragma unroll
        for (int i = 0; i < QS; ++i) {
            const int r = qch + kOsQK * i;
            if (r < M) {
                if (!wide) ud[(size_t)r * N + qc] = u[i];
                else ud[(size_t)qc * M + r] = u[i];
            }
        }
}

// ---- the per-matrix solver -----------------------------------------------

template <int NP, int TPR, int NW, int CHH>
__device__ __forceinline__ void os_solve(const Item& itm, const Pools& P, double* am, double* vb, double* refl, OsShared& sh) {
    using C = OsCfg<NP, TPR, NW, CHH>;
    constexpr bool WFREE = NP >= kOsWFreeMinNP;
    static_assert(!WFREE || (NW == 8 && 3 * NW * C::NL <= os_area_doubles(NP) &&
                             kOsRows * (kOsKC + 1) + kOsKC * (NP + 1) <= os_area_doubles(NP) &&
                             kOsRows * (kOsCKF + 1) <= os_area_doubles(NP)),
                  "the W-free LDS area holds the partial sums and the staged chunks");
    constexpr int NL = C::NL;
    constexpr int NT = C::NT, CH = C::CH, PC = C::P, PP = C::PP, XR = C::XR, LDA = C::LDA;
    constexpr int SC = C::SC, NA = C::NA, LB = C::LB, RGW = C::RGW;
    const int tid = tix(), w = tid >> 6, lane = tid & 63;
    const int rgw = lane / TPR, vch = lane % TPR, rg = w * RGW + rgw;
    const int m0 = itm.m, n0 = itm.n;
    const bool wide = m0 < n0;
    const int M = wide ? n0 : m0, N = wide ? m0 : n0;
    const double* A = P.a + itm.a_off;
    const View ain{A, n0, wide ? 1 : 0};
    lsvd_info_t info = {};
    double sc = 0.0;  // ||A||_F
    struct Clk {
        long long* out;
        long long t;
        __device__ void mark(int ph) {
            if (out && threadIdx.x == 0) {
                const long long now = clock64();
                out[ph] += now - t;
                t = now;
            }
        }
    } clk{P.timing ? P.timing + (size_t)itm.idx * kPhases : nullptr, P.timing && threadIdx.x == 0 ? clock64() : 0};

    // global slot: A's R and reflector tails (kOsRows x NP); when W-free
    // also Y's (NP x NP) and V, row-major (NP x NP)
    double* slot_a = refl;
    double* slot_y = refl + (size_t)kOsRows * NP;
    double* slot_v = slot_y + (size_t)NP * NP;
    (void)slot_y;
    (void)slot_v;

    // ---- A / ||A||_F into the column-resident registers, QR.  W-free
    // sizes up to 96 with M <= 96 take 24 row slots a lane (the pivot
    // buffers' rows from 96 on stay zero) and keep a step's pivot values
    // live; the others 32, reading them again when W-free.
    constexpr bool A24 = WFREE && NP <= 96;
    const bool small_a = A24 && M <= kOsQK * 24;  // uniform
    auto load_qr = [&](auto qs_tag) {
        constexpr int QS = decltype(qs_tag)::value;
        static_assert(NT / kOsQK >= NP, "a column slot per column");
        const int qc = tid >> 2, qch = tid & 3;
        if constexpr (QS < kOsQR)
            for (int e = tid; e < 2 * (kOsRows - kOsQK * QS); e += NT)
                vb[(e / (kOsRows - kOsQK * QS)) * kOsRows + kOsQK * QS + e % (kOsRows - kOsQK * QS)] = 0.0;
        double a[QS];
        double amx = 0.0;
#pragma unroll
        for (int i = 0; i < QS; ++i) {
            const int r = qch + kOsQK * i;
            a[i] = (r < M && qc < N) ? ain(r, qc) : 0.0;
            amx = fmax(amx, fabs(a[i]));
        }
        // ||A||_F without overflow or underflow: the entries are first
        // brought to [-1, 1] by the power of two of the largest (exact, so
        // a matrix and its power-of-two multiples give the same factors);
        // NaN and Inf entries propagate to the results
        const int ex = __builtin_amdgcn_frexp_exp(block_max<NT>(amx, sh.red));
        double ss = 0.0;
#pragma unroll
        for (int i = 0; i < QS; ++i) {
            a[i] = ldexp(a[i], -ex);
            ss = fma(a[i], a[i], ss);
        }
        ss = block_sum<NT>(ss, sh.red);
        const double scr = ss > 0.0 ? sqrt(ss) : 1.0;
        sc = ss > 0.0 ? ldexp(scr, ex) : 0.0;
        const double isc = 1.0 / scr;
#pragma unroll
        for (int i = 0; i < QS; ++i) a[i] *= isc;
        if constexpr (WFREE) os_qr_regs<NT, (QS > 24), QS>(M, N, a, slot_a, NP, vb, sh);  // R and tails to the slot
        else os_qr_regs<NT>(M, N, a, am, LDA, vb, sh);
    };
    if constexpr (A24) {
        if (small_a) load_qr(std::integral_constant<int, 24>{});
        else load_qr(std::integral_constant<int, kOsQR>{});
    } else {
        load_qr(std::integral_constant<int, kOsQR>{});
    }
    __syncthreads();
    clk.mark(kPhQR);
    if constexpr (WFREE) {
        for (int j = tid; j < N; j += NT) sh.tau_a[j] = sh.tau[j];
    } else {  // reflector tails to the global slot (read back when U = Q W is formed)
        for (int e = tid; e < M * NP; e += NT) {
            const int i = e / NP, j = e % NP;
            if (i > j && j < N) refl[(size_t)i * NP + j] = am[i * LDA + j];
        }
    }

    // ---- registers: X = R^T (row r, position g holds R[g][r]), W = I
    double x[XR][CH], wv[WFREE ? 1 : XR][CH];
#pragma unroll
    for (int k = 0; k < XR; ++k) {
        const int r = rg * XR + k;
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int g = vch * CH + c;
            if constexpr (WFREE) {
                x[k][c] = (g <= r && r < N) ? slot_a[(size_t)g * NP + r] : 0.0;
            } else {
                x[k][c] = (g <= r && r < N) ? am[g * LDA + r] : 0.0;
                wv[k][c] = g == r ? 1.0 : 0.0;
            }
        }
    }
    __syncthreads();  // am is reused below

    const double tol = sqrt((double)N) * 0x1p-52;
    const double tol2 = tol * tol;
    // The ordering runs over NE = N rounded up to whole lanes' positions
    // rather than the padded size: the positions from NE on hold zero
    // columns behind the dead pair of the lane that ends at NE - 1, so a
    // sweep takes NE layers.
    const int NE = (N + CH - 1) / CH * CH;
    const int maxG = (60 + 20 * N) * NE;
    int G = 0, quiet = 0, rots = 0;
    // the lane whose last position is NE - 1, and a row's last lane: their
    // straddling pairs are dead; the lanes after them start a segment
    const bool lastv = vch == NE / CH - 1 || vch == TPR - 1, firstv = vch == 0 || vch == NE / CH;
    // lane bits that the reduce-scatter did not split (replicated sums)
    const int rep_mask = (RGW - 1) & ~((SC >= 1 ? (RGW >> 1) : 0) | (SC >= 2 ? (RGW >> 2) : 0));
    const bool rep_lane = (rgw & rep_mask) == 0 && w == 0;
    // Two positions per lane (one pair per layer, its angle formed by the
    // lane itself, replicated over rowgroups and waves): the squared column
    // norms travel with the columns in registers (nr0, nr1 for the lane's
    // positions 0 and 1) and are updated from each rotation, so a layer
    // reduces only g.  The Gram diagonal is formed from the columns again on
    // the first layer of every sweep (an even layer: every position is in a
    // pair) and on the two layers after a rotation whose norm update lost
    // more than kOsRefresh of a column.  The decisions are the same in every
    // wave (the reduced sums are bitwise equal there), so the layer's branch
    // is uniform.
    constexpr bool INC = CH == 2;
    double nr0 = 0.0, nr1 = 0.0;
    bool refresh = false;

    // one layer; ODD: pairs (2j+1, 2j+2) in the lane and (CH-1, next lane's 0);
    // FULL: all three Gram entries from the columns (always, unless INC)
    auto layer = [&](auto odd_tag, const bool full_in) {
        constexpr bool ODD = decltype(odd_tag)::value;
        const bool full = !INC || full_in;
        auto lo = [](int j) { return ODD ? 2 * j + 1 : 2 * j; };
        // partial Gram entries of the lane's pairs over its rows
        double v[PP][3];
#pragma unroll
        for (int j = 0; j < PP; ++j) v[j][0] = v[j][1] = v[j][2] = 0.0;
        // the straddling pair's partner: the next lane's first position (kept
        // for the rotation of X)
        double xnx[ODD ? XR : 1];
#pragma unroll
        for (int k = 0; k < (ODD ? XR : 1); ++k) xnx[k] = ODD ? os_next0(x[k][0]) : 0.0;
        if (full) {
#pragma unroll
            for (int k = 0; k < XR; ++k) {
                const double xn = ODD ? xnx[k] : 0.0;
#pragma unroll
                for (int j = 0; j < PC; ++j) {
                    const double p = x[k][lo(j)];
                    const double q = (ODD && j == PC - 1) ? xn : x[k][lo(j) + 1];
                    v[j][0] = fma(p, p, v[j][0]);
                    v[j][1] = fma(q, q, v[j][1]);
                    v[j][2] = fma(p, q, v[j][2]);
                }
            }
        } else {  // g only: two chains over the rows (even and odd k)
            double g1[PP];
#pragma unroll
            for (int j = 0; j < PP; ++j) g1[j] = 0.0;
#pragma unroll
            for (int k = 0; k < XR; ++k) {
                const double xn = ODD ? xnx[k] : 0.0;
#pragma unroll
                for (int j = 0; j < PC; ++j) {
                    const double p = x[k][lo(j)];
                    const double q = (ODD && j == PC - 1) ? xn : x[k][lo(j) + 1];
                    if (k & 1) g1[j] = fma(p, q, g1[j]);
                    else v[j][2] = fma(p, q, v[j][2]);
                }
            }
#pragma unroll
            for (int j = 0; j < PC; ++j) v[j][2] += g1[j];
        }
        // reduce-scatter over lane bit 5, then bit 4: pair index owned
        int kid = 0;
        if constexpr (SC >= 1) {
#pragma unroll
            for (int j = 0; j < PP / 2; ++j)
#pragma unroll
                for (int t = 0; t < 3; ++t) {
                    double a = v[j][t], b = v[j + PP / 2][t];
                    xch32(a, b);
                    v[j][t] = a + b;
                }
            kid += (lane >> 5) ? PP / 2 : 0;
        }
        if constexpr (SC >= 2) {
#pragma unroll
            for (int j = 0; j < PP / 4; ++j)
#pragma unroll
                for (int t = 0; t < 3; ++t) {
                    double a = v[j][t], b = v[j + PP / 4][t];
                    xch16(a, b);
                    v[j][t] = a + b;
                }
            kid += ((lane >> 4) & 1) ? PP / 4 : 0;
        }
        // butterflies over the remaining rowgroup bits
#pragma unroll
        for (int j = 0; j < NA; ++j)
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                if (t < 2 && !full) continue;  // uniform
                double s = v[j][t];
                if constexpr (SC < 1 && LB >= 1) s = bfly32(s);
                if constexpr (SC < 2 && LB >= 2) s = bfly16(s);
                if constexpr (LB >= 3) s = bfly8(s);
                v[j][t] = s;
            }
        // across the waves (LDS, one barrier; the buffer alternates by layer)
        if constexpr (NW > 1) {
            // planes [wave][entry][pair]: a wave's lanes touch consecutive pairs
            double* pb = am + (size_t)(ODD ? 1 : 0) * NW * 3 * C::S;
#pragma unroll
            for (int j = 0; j < NA; ++j) {
                const int pr = vch * PC + kid + j;
                if (kid + j < PC) {
                    double* d = pb + (size_t)w * 3 * C::S + pr;
                    if (full) {
                        d[0] = v[j][0];
                        d[C::S] = v[j][1];
                    }
                    d[2 * C::S] = v[j][2];
                }
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < NA; ++j) {
                const int pr = vch * PC + (kid + j < PC ? kid + j : 0);
                // the waves' partials summed as a tree (the same order in
                // every wave)
                double s0 = 0.0, s1 = 0.0;
                if (full) {
                    s0 = os_tree_sum<NW>(pb + pr, 3 * C::S);
                    s1 = os_tree_sum<NW>(pb + C::S + pr, 3 * C::S);
                }
                const double s2 = os_tree_sum<NW>(pb + 2 * C::S + pr, 3 * C::S);
                v[j][0] = s0;
                v[j][1] = s1;
                v[j][2] = s2;
            }
        }
        if constexpr (INC) {  // the Gram diagonal from the norms the columns carry
            const double nb = ODD ? os_next0(nr0) : nr1;
            if (!full) {
                v[0][0] = ODD ? nr1 : nr0;
                v[0][1] = nb;
            }
        }
        // angles of the owned pairs
        double cs[PP][2];
        int nrot = 0;
        bool any = false, lost = false;
        double nlo = 0.0, nhi = 0.0;  // INC: the pair's squared norms after the layer (low, high position)
#pragma unroll
        for (int j = 0; j < NA; ++j) {
            const int jk = kid + j;
            const bool boundary_dead = ODD && lastv && jk == PC - 1;  // no partner across the segment end
            const bool valid = jk < PC && !boundary_dead;
            const double a = v[j][0], b = v[j][1], g = v[j][2];
            const bool rot = valid && g * g > tol2 * (a * b);
            double c = 1.0, s = 0.0;
            if constexpr (INC) {
                double t = 0.0;
                if (rot) os_angle<true>(a, b, g, c, s, &t);
                const double tg = rot ? t * g : 0.0;
                nlo = boundary_dead ? a : b + tg;
                nhi = a - tg;
                lost = rot && (nhi < kOsRefresh * a || nlo < kOsRefresh * b);
            } else {
                if (rot) os_angle(a, b, g, c, s);
            }
            if (boundary_dead) {
                c = 0.0;
                s = 1.0;
            }
            cs[j][0] = c;
            cs[j][1] = s;
            nrot += (rot && rep_lane) ? 1 : 0;
            any = any || rot;
        }
        if constexpr (INC) {  // the norms trade with the columns
            if constexpr (ODD) {
                const double ph = os_prev0(nhi);  // the previous lane's pair's high position is this lane's 0
                nr1 = nlo;
                if (!firstv) nr0 = ph;
            } else {
                nr0 = nlo;
                nr1 = nhi;
            }
            refresh = __any(lost);
        }
        // gather the (c, s) of all the lane's pairs (reverse of the scatter)
        if constexpr (SC >= 2) {
#pragma unroll
            for (int j = PP / 4 - 1; j >= 0; --j)
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    double a = cs[j][t], b = cs[j][t];
                    xch16(a, b);
                    cs[j][t] = a;
                    cs[j + PP / 4][t] = b;
                }
        }
        if constexpr (SC >= 1) {
#pragma unroll
            for (int j = PP / 2 - 1; j >= 0; --j)
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    double a = cs[j][t], b = cs[j][t];
                    xch32(a, b);
                    cs[j][t] = a;
                    cs[j + PP / 2][t] = b;
                }
        }
        // rotate and trade: the pair's low position takes s p + c q, the high c p - s q
        double cpv = 1.0, spv = 0.0;
        if constexpr (ODD) {  // the previous lane's straddling pair
            cpv = os_from_prev(cs[PC - 1][0], 0.0);
            spv = os_from_prev(cs[PC - 1][1], 1.0);
            if (firstv) spv = -spv;  // row start: (0, 1) of the dead pair -> keep q
        }
        auto rot_rows = [&](double (*m)[CH], auto is_x_tag) {
#pragma unroll
            for (int k = 0; k < XR; ++k) {
                double mn = 0.0, mp = 0.0;
                if constexpr (ODD) {  // the neighbours' boundary columns, before this row changes
                    if constexpr (decltype(is_x_tag)::value) mn = xnx[k];
                    else mn = os_next0(m[k][0]);
                    mp = os_prev0(m[k][CH - 1]);
                }
                double nx[CH];
#pragma unroll
                for (int j = 0; j < PC; ++j) {
                    const int l = lo(j);
                    const double c = cs[j][0], s = cs[j][1];
                    const double p = m[k][l];
                    if (ODD && j == PC - 1) {
                        // the addend is this column scaled in place, so the
                        // accumulating FMA leaves the result in the
                        // column's own register (no copies at the loop's
                        // back edge; as nx[0] below)
                        nx[l] = fma(c, mn, s * p);
                    } else {
                        const double q = m[k][l + 1];
                        nx[l] = fma(s, p, c * q);
                        nx[l + 1] = fma(c, p, -s * q);
                    }
                }
                if constexpr (ODD) nx[0] = fma(cpv, mp, -spv * m[k][0]);
#pragma unroll
                for (int c = 0; c < CH; ++c) m[k][c] = nx[c];
            }
        };
        rot_rows(x, std::true_type{});
        if constexpr (!WFREE) rot_rows(wv, std::false_type{});
        any = __any(any);
        rots += nrot;
        return any;
    };

    if (N > 1) {
        int since = 0, full_left = 0;  // layers since the sweep's first; full layers still owed
        while (quiet < NE && G < maxG) {
            const bool a0 = layer(std::integral_constant<bool, false>{}, since == 0 || full_left > 0);
            full_left = refresh ? 2 : (full_left > 0 ? full_left - 1 : 0);
            ++G;
            quiet = a0 ? 0 : quiet + 1;
            if (quiet >= NE) break;
            const bool a1 = layer(std::integral_constant<bool, true>{}, full_left > 0);
            full_left = refresh ? 2 : (full_left > 0 ? full_left - 1 : 0);
            ++G;
            quiet = a1 ? 0 : quiet + 1;
            since += 2;
            if (since >= NE) since = 0;
        }
    }
    __syncthreads();  // every wave's last partial-sum reads are done before am is reused
    clk.mark(kPhJacobi);
    info.ns_iters = 0;
    info.jacobi_sweeps = (G - quiet + NE - 1) / NE;
    info.polar_ok = 1;

    // ---- column norms -> sigma (normalised), rank
    {
        double cn[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < XR; ++k) s = fma(x[k][c], x[k][c], s);
            if constexpr (LB >= 1) s = bfly32(s);
            if constexpr (LB >= 2) s = bfly16(s);
            if constexpr (LB >= 3) s = bfly8(s);
            cn[c] = s;
        }
        if (rgw == 0)
#pragma unroll
            for (int c = 0; c < CH; ++c) am[w * NL + vch * CH + c] = cn[c];
        __syncthreads();
        for (int g = tid; g < NE; g += NT) {
            double s = 0.0;
            for (int ww = 0; ww < NW; ++ww) s += am[ww * NL + g];
            sh.sig[os_index_at(NE, G, g)] = sqrt(s);
        }
        __syncthreads();
    }
    for (int i = tid; i < N; i += NT) {  // rank: descending, NaN last, ties by index
        const double di = isnan(sh.sig[i]) ? -INFINITY : sh.sig[i];
        int r = 0;
        for (int j = 0; j < N; ++j) {
            const double dj = isnan(sh.sig[j]) ? -INFINITY : sh.sig[j];
            r += (dj > di) || (dj == di && j < i);
        }
        sh.rank[i] = r;
    }
    if (tid == 0) sh.flag = 0;
    __syncthreads();

    if constexpr (!WFREE) {
    // ---- V = B diag(sigma)^-1 in sorted column order (am: N x LDA)
#pragma unroll
    for (int k = 0; k < XR; ++k) {
        const int r = rg * XR + k;
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int g = vch * CH + c;
            const int idx = g < NE ? os_index_at(NE, G, g) : NE;
            if (r < N && idx < N) {
                // columns whose norm is below kOsTiny (the input is scaled to
                // ||A||_F = 1) are zero to the rotation test (their Gram
                // products underflow), so their direction is not orthogonal to
                // working precision: they are completed instead
                const double sg = sh.sig[idx];
                const bool tiny = !(sg > kOsTiny) && !isnan(sg);
                am[r * LDA + sh.rank[idx]] = tiny ? 0.0 : x[k][c] / sg;
                if (tiny) sh.flag = 1;
            }
        }
    }
    __syncthreads();
    if (sh.flag) {  // zero (or underflowing) columns of B: complete V (the reference's Gram-Schmidt rule)
        gram_schmidt_g<NT>(N, N, View{am, LDA, 0}, am, LDA, sh.red);
        info.repaired = 1;
        __syncthreads();
    }
    {
        double* sdst = P.s + itm.s_off;
        for (int i = tid; i < N; i += NT) sdst[sh.rank[i]] = sh.sig[i] * sc;
        double* vdst = wide ? P.u + itm.u_off : P.vt + itm.vt_off;
        for (int e = tid; e < NP * NP; e += NT) {
            const int r = e / NP, j = e % NP;  // V[r][j]
            if (r < N && j < N) {
                if (!wide) vdst[(size_t)j * N + r] = am[r * LDA + j];
                else vdst[(size_t)r * N + j] = am[r * LDA + j];
            }
        }
    }
    __syncthreads();
    clk.mark(kPhU);

    // ---- U = Q [W; 0] in sorted column order (W: am's first N rows)
#pragma unroll
    for (int k = 0; k < XR; ++k) {
        const int r = rg * XR + k;
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int g = vch * CH + c;
            const int idx = g < NE ? os_index_at(NE, G, g) : NE;
            if (r < N && idx < N) am[r * LDA + sh.rank[idx]] = wv[k][c];
        }
    }
    __syncthreads();
    {
        double u[kOsQR];
        const int qc = tid >> 2, qch = tid & 3;
#pragma unroll
        for (int i = 0; i < kOsQR; ++i) {
            const int r = qch + kOsQK * i;
            u[i] = (r < N && qc < N) ? am[(size_t)r * LDA + qc] : 0.0;
        }
        os_apply_refl<NT, NP, false>(M, N, u, refl, NP, sh.tau, am);  // all reflectors staged at once (am: 128 x LDA)
        os_write_u(u, M, N, wide ? P.vt + itm.vt_off : P.u + itm.u_off, wide);
    }
    } else {
    // ---- V = B diag(sigma)^-1, sorted, row-major in the slot (then out)
#pragma unroll
    for (int k = 0; k < XR; ++k) {
        const int r = rg * XR + k;
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int g = vch * CH + c;
            const int idx = g < NE ? os_index_at(NE, G, g) : NE;
            if (r < N && idx < N) {
                const double sg = sh.sig[idx];  // below kOsTiny: completed as above
                const bool tiny = !(sg > kOsTiny) && !isnan(sg);
                slot_v[(size_t)r * NP + sh.rank[idx]] = tiny ? 0.0 : x[k][c] / sg;
                if (tiny) sh.flag = 1;
            }
        }
    }
    for (int i = tid; i < N; i += NT) sh.tw[sh.rank[i]] = sh.sig[i];  // sigma in sorted order
    __syncthreads();
    if (sh.flag) {
        gram_schmidt_g<NT>(N, N, View{slot_v, NP, 0}, slot_v, NP, sh.red);
        info.repaired = 1;
        __syncthreads();
    }
    {
        double* sdst = P.s + itm.s_off;
        for (int i = tid; i < N; i += NT) sdst[sh.rank[i]] = sh.sig[i] * sc;
        double* vdst = wide ? P.u + itm.u_off : P.vt + itm.vt_off;
        for (int e = tid; e < NP * NP; e += NT) {
            const int r = e / NP, j = e % NP;  // V[r][j]
            if (r < N && j < N) {
                const double v = slot_v[(size_t)r * NP + j];
                if (!wide) vdst[(size_t)j * N + r] = v;
                else vdst[(size_t)r * N + j] = v;
            }
        }
    }
    clk.mark(kPhU);

    // ---- Y = R V diag(sigma)^-1 into the column-resident registers, k in
    // chunks of kOsKC through LDS (R's columns k0.. and V's rows k0..)
    // Y has N <= NP rows: its QR and Q_Y take NP / 4 row slots a lane (the
    // pivot buffers' rows from NP on stay zero)
    constexpr int QY = NP / kOsQK;
    static_assert(QY % 4 == 0 && QY <= kOsQR, "row slots of Y");
    for (int e = tid; e < 2 * (kOsRows - NP); e += NT) vb[(e / (kOsRows - NP)) * kOsRows + NP + e % (kOsRows - NP)] = 0.0;
    const int qc = tid >> 2, qch = tid & 3;
    double y[QY];
#pragma unroll
    for (int i = 0; i < QY; ++i) y[i] = 0.0;
    double* rc = am;                           // NP x (kOsKC + 1)
    double* vc = am + kOsRows * (kOsKC + 1);   // kOsKC x (NP + 1)
    for (int k0 = 0; k0 < N; k0 += kOsKC) {
        __syncthreads();
        for (int e = tid; e < NP * kOsKC; e += NT) {
            const int r = e / kOsKC, kk = e % kOsKC, k = k0 + kk;
            rc[r * (kOsKC + 1) + kk] = (r <= k && k < N) ? slot_a[(size_t)r * NP + k] : 0.0;
        }
        for (int e = tid; e < kOsKC * NP; e += NT) {
            const int kk = e / NP, cc = e % NP, k = k0 + kk;
            vc[kk * (NP + 1) + cc] = (k < N && cc < N) ? slot_v[(size_t)k * NP + cc] : 0.0;
        }
        __syncthreads();
        if (qc < N) {
#pragma unroll
            for (int kk = 0; kk < kOsKC; ++kk) {
                const double v = vc[kk * (NP + 1) + qc];
#pragma unroll
                for (int i = 0; i < QY; ++i) y[i] = fma(rc[(qch + kOsQK * i) * (kOsKC + 1) + kk], v, y[i]);
            }
        }
    }
    {
        const double sg = qc < N ? sh.tw[qc] : 0.0;
        const double is = (sg > kOsTiny) ? 1.0 / sg : 0.0;  // a completed V column: Y's column is zero
#pragma unroll
        for (int i = 0; i < QY; ++i) y[i] *= is;
    }
    // ---- Y = Q_Y R_Y (Y's reflectors and R_Y to the slot), U = Q_A [Q_Y D; 0]
    os_qr_regs<NT, kOsYTight<NP>, QY>(N, N, y, slot_y, NP, vb, sh);
    __syncthreads();
    {
        double uy[QY];
        const double d = qc < N ? (slot_y[(size_t)qc * NP + qc] < 0.0 ? -1.0 : 1.0) : 0.0;
#pragma unroll
        for (int i = 0; i < QY; ++i) uy[i] = (qch + kOsQK * i == qc) ? d : 0.0;
        os_apply_refl<NT, kOsCKF, kOsYTight<NP>, QY>(N, N, uy, slot_y, NP, sh.tau, am);
        // Q_A on M rows: as many row slots as A's QR took
        auto apply_a = [&](auto qs_tag) {
            constexpr int QS = decltype(qs_tag)::value;
            static_assert(QS >= QY, "Q_Y D fits the slots");
            double u[QS];
#pragma unroll
            for (int i = 0; i < QS; ++i) u[i] = i < QY ? uy[i] : 0.0;
            os_apply_refl<NT, kOsCKF, (QS > 24), QS>(M, N, u, slot_a, NP, sh.tau_a, am);
            os_write_u(u, M, N, wide ? P.vt + itm.vt_off : P.u + itm.u_off, wide);
        };
        if constexpr (A24) {
            if (small_a) apply_a(std::integral_constant<int, 24>{});
            else apply_a(std::integral_constant<int, kOsQR>{});
        } else {
            apply_a(std::integral_constant<int, kOsQR>{});
        }
    }
    }
    clk.mark(kPhOut);
    rots = (int)block_sum<NT>((double)rots, sh.red);
    info.rotations = rots;
    if (tid == 0 && P.info) {
        info.path = itm.mode == kModeMicroFallback ? 2 : 4;
        info.micro_fallback = itm.mode == kModeMicroFallback ? 1 : 0;
        P.info[itm.idx] = info;
    }
    __syncthreads();
}

// ---- kernels ---------------------------------------------------------------

template <int NP, int TPR = OsDefault<NP>::TPR, int NW = OsDefault<NP>::NW, int CHH = OsDefault<NP>::CH>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(os_min_waves(NP)))) void os_svd_kernel(const Item* __restrict__ items, const int* count_ptr,
                                                         int* counter, Pools pools, double* gws, int64_t gws_stride) {
    using C = OsCfg<NP, TPR, NW, CHH>;
    extern __shared__ double dyn_lds[];
    __shared__ OsShared sh;
    const int count = *count_ptr;
    double* refl = gws + (size_t)blockIdx.x * gws_stride;
    // LDS: the QR's two reflector buffers (2 x 128 doubles), then the matrix
    // area (os_area_doubles: R, W and V below padded size 80; the sweeps'
    // partial sums and staged chunks)
    double* vb = dyn_lds;
    double* am = dyn_lds + 2 * kOsRows;
    for (;;) {
        if (threadIdx.x == 0) sh.item = atomicAdd(counter, 1);
        __syncthreads();
        const int it = sh.item;
        __syncthreads();
        if (it >= count) break;
        const Item itm = items[it];
        os_solve<NP, TPR, NW, CHH>(itm, pools, am, vb, refl, sh);
    }
}

template <int NP>
using OsDefCfg = OsCfg<NP, OsDefault<NP>::TPR, OsDefault<NP>::NW, OsDefault<NP>::CH>;

// One persistent launch over the padded sizes 80..128, which share the
// launch shape (64 lanes per row, 8 waves, 2 positions per lane): the queue
// holds the items ordered by padded size, largest first; count_ptr[1 + j]
// is the end of size class j (128, 112, 96, 80).  Each workgroup drains the
// classes in that order with the solver instance of each size; an index
// fetched past a class's end is carried into the next class.
constexpr int kOsMultiNT = 512;
static_assert(OsDefCfg<80>::NT == kOsMultiNT && OsDefCfg<96>::NT == kOsMultiNT &&
                  OsDefCfg<112>::NT == kOsMultiNT && OsDefCfg<128>::NT == kOsMultiNT,
              "one block size for the multi-size launch");

template <int NP>
__device__ __forceinline__ void os_multi_drain(const Item* items, int end, int* counter, const Pools& pools,
                                               double* am, double* vb, double* refl, OsShared& sh,
                                               int* next_item) {
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
        os_solve<NP, OsDefault<NP>::TPR, OsDefault<NP>::NW, OsDefault<NP>::CH>(itm, pools, am, vb, refl, sh);
        __syncthreads();
    }
}

__global__ __launch_bounds__(kOsMultiNT) __attribute__((amdgpu_waves_per_eu(os_min_waves(kOsMultiMinNP)))) void os_svd_multi_kernel(const Item* __restrict__ items,
                                                                  const int* count_ptr, int* counter, Pools pools,
                                                                  double* gws, int64_t gws_stride) {
    extern __shared__ double dyn_lds[];
    __shared__ OsShared sh;
    __shared__ int next_item;
    double* refl = gws + (size_t)blockIdx.x * gws_stride;
    double* vb = dyn_lds;
    double* am = dyn_lds + 2 * kOsRows;
    if (threadIdx.x == 0) next_item = -1;
    __syncthreads();
    os_multi_drain<128>(items, count_ptr[1], counter, pools, am, vb, refl, sh, &next_item);
    os_multi_drain<112>(items, count_ptr[2], counter, pools, am, vb, refl, sh, &next_item);
    os_multi_drain<96>(items, count_ptr[3], counter, pools, am, vb, refl, sh, &next_item);
    os_multi_drain<80>(items, count_ptr[4], counter, pools, am, vb, refl, sh, &next_item);
}

hipError_t launch_onesided_multi(const SmallLaunch& L, hipStream_t st) {
    if (L.lds_bytes > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&os_svd_multi_kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)L.lds_bytes);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(os_svd_multi_kernel, dim3(L.grid), dim3(kOsMultiNT), L.lds_bytes, st, L.items, L.count_ptr,
                       L.counter, L.pools, L.gws, L.gws_stride);
    return hipGetLastError();
}

int onesided_multi_blocks_per_cu(size_t lds) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, os_svd_multi_kernel, kOsMultiNT, lds) != hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

#define LSVD_OS_SIZES(X) X(16) X(32) X(48) X(64) X(80) X(96) X(112) X(128)

int os_threads(int np) {
    switch (np) {
#define X(np_) \
    case np_:  \
        return OsDefCfg<np_>::NT;
        LSVD_OS_SIZES(X)
#undef X
    }
    return 0;
}

size_t os_lds_bytes(int np, int M) {
    (void)M;  // the register layout's row slots cover 128 rows whatever M is
    return sizeof(double) * (2 * (size_t)kOsRows + (size_t)os_area_doubles(np));
}

int64_t os_slot_doubles(int np, int M) {
    if (np >= kOsWFreeMinNP)  // A's R and tails (128 rows), Y's, V
        return ((int64_t)kOsRows * np + 2 * (int64_t)np * np + 31) / 32 * 32;
    const int mp = (M + 15) & ~15;
    const int rows = mp > np ? mp : np;
    return ((int64_t)rows * np + 31) / 32 * 32;
}

template <int NP>
static hipError_t os_launch(const SmallLaunch& L, hipStream_t st) {
    constexpr int NT = OsDefCfg<NP>::NT;
    if (L.lds_bytes > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&os_svd_kernel<NP>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)L.lds_bytes);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((os_svd_kernel<NP>), dim3(L.grid), dim3(NT), L.lds_bytes, st, L.items, L.count_ptr, L.counter,
                       L.pools, L.gws, L.gws_stride);
    return hipGetLastError();
}

template <int NP>
static int os_blocks(size_t lds) {
    constexpr int NT = OsDefCfg<NP>::NT;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, os_svd_kernel<NP>, NT, lds) != hipSuccess) return 1;
    return nb > 0 ? nb : 1;
}

hipError_t launch_onesided(int np, const SmallLaunch& L, hipStream_t st) {
    switch (np) {
#define X(np_) \
    case np_:  \
        return os_launch<np_>(L, st);
        LSVD_OS_SIZES(X)
#undef X
    }
    return hipErrorInvalidValue;
}

int onesided_blocks_per_cu(int np, size_t lds) {
    switch (np) {
#define X(np_) \
    case np_:  \
        return os_blocks<np_>(lds);
        LSVD_OS_SIZES(X)
#undef X
    }
    return 1;
}

}  // namespace lsvd
