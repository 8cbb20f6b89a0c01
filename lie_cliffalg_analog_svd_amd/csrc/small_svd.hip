// small_svd.hip -- batched FP64 LieSvdSmall for gfx950.
//
// One workgroup solves one matrix at a time and pulls the next one from a
// cost-sorted queue (persistent kernel).  Per matrix (reference
// src/lie_svd_small.rs:296-360):
//   tall/wide  -> Householder QR front end (the reference's qr_reduce is
//                 MGS; see solve_item for why Householder is used here)
//   square R   -> Newton-Schulz polar X <- X(3I - X^T X)/2      (:158-183)
//              -> P = sym(Q^T R)                                 (:305-309)
//              -> two-sided Jacobi on P, round-robin layers of
//                 disjoint pairs, Rutishauser tangent            (:212-285)
//              -> sort descending, U_r = Q V, repair if needed   (:313-320)
//   tall       -> U = Q_qr [U_r; 0]
//
// Working matrices live in an arena of cap x ld doubles with an odd
// leading dimension (ld = N | 1) against LDS bank conflicts.  The arena is
// dynamic LDS for small N and a per-workgroup global slot (L2 resident)
// otherwise; the same code serves both through flat pointers.
#include <hip/hip_runtime.h>
#include <math.h>

#include "lsvd_device.hpp"
#include "lsvd_kernels.hpp"
#include "lsvd_small_common.hpp"

namespace lsvd {

// The square matrix the solve runs on: the input itself, or the R factor
// of the front-end QR (upper triangle of the compact form in global memory).
struct AsqView {
    UpperView r;
    View a;
    bool tall;
    __device__ __forceinline__ double operator()(int i, int j) const { return tall ? r(i, j) : a(i, j); }
};

template <int NT, int MAXN>
struct SmallShared {
    double red[NT / kWave + 2];
    double diag[MAXN];
    double tau[MAXN];
    double vbuf[MAXN > 128 ? MAXN : 128];
    double rc[MAXN / 2 + 1], rs[MAXN / 2 + 1], rt[MAXN / 2 + 1];
    int rp[MAXN / 2 + 1], rq[MAXN / 2 + 1], rf[MAXN / 2 + 1];
    int rank[MAXN];
    int item;
};

// ||.||_F of an N x N view.
template <int NT, class V>
__device__ double frob_view(int R, int Cc, const V& v, double* red) {
    double s = 0.0;
    for (int e = threadIdx.x; e < R * Cc; e += NT) {
        const double x = v(e / Cc, e % Cc);
        s += x * x;
    }
    return sqrt(block_sum<NT>(s, red));
}

// ||Q^T Q - I||_F for Q (N x N at q, ld)  (src/lie_svd_small.rs:146-150).
template <int NT>
__device__ double orth_err(int N, const double* q, int ld, double* red) {
    View qt{q, ld, 1}, qn{q, ld, 0};
    double s = 0.0;
    wg_gemm<NT>(N, N, N, qt, qn, [&](int i, int j, double acc) {
        const double d = acc - (i == j ? 1.0 : 0.0);
        s += d * d;
    });
    return sqrt(block_sum<NT>(s, red));
}

// Thin wrappers over the shared building blocks (lsvd_small_common.hpp).
template <int NT, int MAXN, class V>
__device__ void gram_schmidt(int N, const V& a, double* out, int ld, SmallShared<NT, MAXN>& sh) {
    gram_schmidt_g<NT>(N, N, a, out, ld, sh.red);
}

template <int NT, int MAXN>
__device__ void householder_qr(int M, int N, double* a, int ld, SmallShared<NT, MAXN>& sh) {
    householder_qr_g<NT>(M, N, a, ld, sh.red, sh.tau, sh.diag);
}

template <int NT, int MAXN>
__device__ void apply_q(int M, int N, const double* qs, double* y, int ld, SmallShared<NT, MAXN>& sh) {
    apply_q_g<NT>(M, N, qs, y, ld, sh.tau, sh.diag);
}

// Parallel two-sided Jacobi on the symmetric P (N x N at t, ld) with
// eigenvectors accumulated in v.  Each layer of the round-robin schedule
// holds disjoint pairs, so its rotations commute and are applied together:
// every 2x2 block (pair a rows, pair b cols) is rotated once in place.
template <int NT, int MAXN>
__device__ void jacobi_eigh(int N, double* t, double* v, int ld, SmallShared<NT, MAXN>& sh, int& sweeps_out,
                            int& rot_out) {
    const double ref = fmax(frob_view<NT>(N, N, View{t, ld, 0}, sh.red), 1e-300);
    const int max_sweeps = 60 + 20 * N;
    const int slots = (N + (N & 1)) / 2;
    const int layers = N > 1 ? rr_layers(N) : 0;
    const int nblk = slots * (slots + 1) / 2;
    int sweeps = 0, rots = 0;
    for (int sw = 0; sw < max_sweeps; ++sw) {
        double off = 0.0;
        for (int e = threadIdx.x; e < N * N; e += NT) {
            const int p = e / N, q = e % N;
            if (q > p) off += t[(size_t)p * ld + q] * t[(size_t)p * ld + q];
        }
        off = block_sum<NT>(off, sh.red);
        if (!(sqrt(off) >= kJacobiTol * ref)) break;  // also stops on NaN
        ++sweeps;
        for (int layer = 0; layer < layers; ++layer) {
            for (int k = threadIdx.x; k < slots; k += NT) {
                int p, q;
                rr_pair(N, layer, k, p, q);
                double c = 1.0, s = 0.0, tt = 0.0;
                int rot = 0;
                if (q >= 0) {
                    const double apq = t[(size_t)p * ld + q];
                    if (fabs(apq) >= 1e-300) {  // the reference skips |a_pq| < 1e-300
                        const double app = t[(size_t)p * ld + p], aqq = t[(size_t)q * ld + q];
                        const double th = (aqq - app) / (2.0 * apq);
                        tt = th >= 0.0 ? 1.0 / (th + sqrt(th * th + 1.0)) : -1.0 / (-th + sqrt(th * th + 1.0));
                        c = 1.0 / sqrt(tt * tt + 1.0);
                        s = tt * c;
                        rot = 1;
                        ++rots;
                    }
                }
                sh.rp[k] = p;
                sh.rq[k] = q;
                sh.rf[k] = rot;
                sh.rc[k] = c;
                sh.rs[k] = s;
                sh.rt[k] = tt;
            }
            __syncthreads();
            for (int b = threadIdx.x; b < nblk; b += NT) {
                // decode b -> (ka <= kb)
                int kb = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
                while (kb * (kb + 1) / 2 > b) --kb;
                while ((kb + 1) * (kb + 2) / 2 <= b) ++kb;
                const int ka = b - kb * (kb + 1) / 2;
                const int pa = sh.rp[ka], qa = sh.rq[ka], pb = sh.rp[kb], qb = sh.rq[kb];
                if (ka == kb) {
                    if (sh.rf[ka]) {  // (:244-249): exact diagonal update, zeroed pair
                        const double apq = t[(size_t)pa * ld + qa];
                        const double tt = sh.rt[ka];
                        t[(size_t)pa * ld + pa] -= tt * apq;
                        t[(size_t)qa * ld + qa] += tt * apq;
                        t[(size_t)pa * ld + qa] = 0.0;
                        t[(size_t)qa * ld + pa] = 0.0;
                    }
                    continue;
                }
                if (!sh.rf[ka] && !sh.rf[kb]) continue;  // both identity
                // qa/qb == -1: phantom partner (odd N), that row/col is absent.
                // A skipped pair keeps its second index with c = 1, s = 0.
                const double ca = sh.rc[ka], sa = sh.rs[ka], cb = sh.rc[kb], sb = sh.rs[kb];
                const double b00 = t[(size_t)pa * ld + pb];
                const double b01 = qb >= 0 ? t[(size_t)pa * ld + qb] : 0.0;
                const double b10 = qa >= 0 ? t[(size_t)qa * ld + pb] : 0.0;
                const double b11 = (qa >= 0 && qb >= 0) ? t[(size_t)qa * ld + qb] : 0.0;
                const double l00 = ca * b00 - sa * b10, l01 = ca * b01 - sa * b11;
                const double l10 = sa * b00 + ca * b10, l11 = sa * b01 + ca * b11;
                const double r00 = cb * l00 - sb * l01, r01 = sb * l00 + cb * l01;
                const double r10 = cb * l10 - sb * l11, r11 = sb * l10 + cb * l11;
                t[(size_t)pa * ld + pb] = r00;
                t[(size_t)pb * ld + pa] = r00;
                if (qb >= 0) {
                    t[(size_t)pa * ld + qb] = r01;
                    t[(size_t)qb * ld + pa] = r01;
                }
                if (qa >= 0) {
                    t[(size_t)qa * ld + pb] = r10;
                    t[(size_t)pb * ld + qa] = r10;
                }
                if (qa >= 0 && qb >= 0) {
                    t[(size_t)qa * ld + qb] = r11;
                    t[(size_t)qb * ld + qa] = r11;
                }
            }
            for (int e = threadIdx.x; e < N * slots; e += NT) {
                const int r = e / slots, k = e % slots;
                if (!sh.rf[k]) continue;
                const int p = sh.rp[k], q = sh.rq[k];
                const double c = sh.rc[k], s = sh.rs[k];
                const double vp = v[(size_t)r * ld + p], vq = v[(size_t)r * ld + q];
                v[(size_t)r * ld + p] = c * vp - s * vq;
                v[(size_t)r * ld + q] = s * vp + c * vq;
            }
            __syncthreads();
        }
    }
    sweeps_out = sweeps;
    rot_out = rots;
}

template <int NT, int MAXN>
__device__ void sort_desc(int N, SmallShared<NT, MAXN>& sh) {
    for (int i = threadIdx.x; i < N; i += NT) {
        const double di = isnan(sh.diag[i]) ? -INFINITY : sh.diag[i];
        int r = 0;
        for (int j = 0; j < N; ++j) {
            const double dj = isnan(sh.diag[j]) ? -INFINITY : sh.diag[j];
            r += (dj > di) || (dj == di && j < i);
        }
        sh.rank[i] = r;
    }
    __syncthreads();
}

// Phase clock: thread 0 accumulates clock64() deltas into Pools::timing.
struct PhaseClock {
    long long* out;
    long long t;
    __device__ PhaseClock(long long* o) : out(o), t(o && threadIdx.x == 0 ? clock64() : 0) {}
    __device__ __forceinline__ void mark(int ph) {
        if (out && threadIdx.x == 0) {
            const long long now = clock64();
            out[ph] += now - t;
            t = now;
        }
    }
};

template <int NT, int MAXN>
__device__ void solve_item(const Item& itm, const Pools& P, double* arena, SmallShared<NT, MAXN>& sh) {
    PhaseClock clk(P.timing ? P.timing + (size_t)itm.idx * kPhases : nullptr);
    const int m0 = itm.m, n0 = itm.n;
    const bool wide = m0 < n0;
    const int M = wide ? n0 : m0, N = wide ? m0 : n0;
    const int ld = N | 1;
    const int cap = (3 * N > M + N) ? 3 * N : M + N;
    double* X = arena;
    double* W = arena + (size_t)N * ld;
    double* T = arena + (size_t)(cap - N) * ld;
    const double* A = P.a + itm.a_off;
    const View ain{A, n0, wide ? 1 : 0};
    double* qs = wide ? P.vt + itm.vt_off : P.u + itm.u_off;
    const double rootn = sqrt((double)N);
    lsvd_info_t info = {};

    if (itm.mode == kModeEigh) {
        // eigh_jacobi_full on the symmetric input (src/lie_svd_small.rs:212-285)
        for (int e = threadIdx.x; e < N * N; e += NT) {
            const int i = e / N, j = e % N;
            T[(size_t)i * ld + j] = A[(size_t)i * N + j];
            W[(size_t)i * ld + j] = (i == j) ? 1.0 : 0.0;
        }
        __syncthreads();
        int sweeps, rots;
        jacobi_eigh<NT, MAXN>(N, T, W, ld, sh, sweeps, rots);
        for (int i = threadIdx.x; i < N; i += NT) sh.diag[i] = T[(size_t)i * ld + i];
        __syncthreads();
        sort_desc<NT, MAXN>(N, sh);
        double* vout = P.u + itm.u_off;
        double* eout = P.s + itm.s_off;
        for (int i = threadIdx.x; i < N; i += NT) eout[sh.rank[i]] = sh.diag[i];
        for (int e = threadIdx.x; e < N * N; e += NT) {
            const int r = e / N, i = e % N;
            vout[(size_t)r * N + sh.rank[i]] = W[(size_t)r * ld + i];
        }
        rots = (int)block_sum<NT>((double)rots, sh.red);
        if (threadIdx.x == 0 && P.info) {
            info.jacobi_sweeps = sweeps;
            info.rotations = rots;
            P.info[itm.idx] = info;
        }
        __syncthreads();
        return;
    }

    // ---- front end: square input is used in place; tall/wide go through QR.
    // The reference's qr_reduce is MGS (src/lie_svd_small.rs:373-418); MGS
    // needs O(N^2) dependent block reductions, Householder needs O(N), and
    // both give an orthonormal Q with Q R = A, so sigma and V agree to
    // rounding.  Difference: for numerically dependent columns the
    // reference zeroes that Q column, whereas Householder keeps Q
    // orthonormal (U stays orthonormal; U S Vt = A either way).
    const bool tall = M > N;
    if (tall) {
        for (int e = threadIdx.x; e < M * N; e += NT) {
            const int i = e / N, j = e % N;
            X[(size_t)i * ld + j] = ain(i, j);
        }
        __syncthreads();
        householder_qr<NT, MAXN>(M, N, X, ld, sh);
        for (int e = threadIdx.x; e < M * N; e += NT) {
            const int i = e / N, j = e % N;
            qs[(size_t)i * N + j] = X[(size_t)i * ld + j];
        }
        __syncthreads();
    }
    const AsqView asv{UpperView{qs, N}, ain, tall};
    clk.mark(kPhQR);

    // ---- Newton-Schulz polar (src/lie_svd_small.rs:158-183)
    double scale = 0.0;
    for (int e = threadIdx.x; e < N * N; e += NT) {
        const double x = asv(e / N, e % N);
        scale += x * x;
    }
    scale = fmax(sqrt(block_sum<NT>(scale, sh.red)), 1e-300);
    for (int e = threadIdx.x; e < N * N; e += NT) {
        const int i = e / N, j = e % N;
        X[(size_t)i * ld + j] = asv(i, j) / scale;
    }
    __syncthreads();
    int polar_ok = 0, iters = 0, done = 0;
    for (int it = 0; it < kPolarMaxIters; ++it) {
        wg_gemm<NT>(N, N, N, View{X, ld, 1}, View{X, ld, 0}, [&](int i, int j, double acc) {
            T[(size_t)i * ld + j] = (i == j ? 3.0 : 0.0) - acc;
        });
        __syncthreads();
        double dsq = 0.0;
        int bad = 0;
        wg_gemm<NT>(N, N, N, View{X, ld, 0}, View{T, ld, 0}, [&](int i, int j, double acc) {
            const double xn = 0.5 * acc;
            W[(size_t)i * ld + j] = xn;
            const double d = xn - X[(size_t)i * ld + j];
            dsq += d * d;
            bad |= !isfinite(xn);
        });
        dsq = block_sum<NT>(dsq, sh.red);
        bad = block_or<NT>(bad, sh.red);
        iters = it + 1;
        const double delta = sqrt(dsq);
        if (bad || delta > 1e6 * rootn) {
            gram_schmidt<NT, MAXN>(N, asv, X, ld, sh);
            polar_ok = 0;
            done = 1;
            break;
        }
        double* tmp = X;
        X = W;
        W = tmp;
        if (delta < kPolarTol * rootn) {
            polar_ok = orth_err<NT>(N, X, ld, sh.red) <= 1e-8 * rootn;
            done = 1;
            break;
        }
    }
    if (!done) polar_ok = orth_err<NT>(N, X, ld, sh.red) <= 1e-8 * rootn;
    info.polar_ok = polar_ok;
    info.ns_iters = iters;
    clk.mark(kPhNS);

    // ---- P = sym(Q^T A)  (:305-309)
    wg_gemm<NT>(N, N, N, View{X, ld, 1}, asv, [&](int i, int j, double acc) { W[(size_t)i * ld + j] = acc; });
    __syncthreads();
    for (int e = threadIdx.x; e < N * N; e += NT) {
        const int i = e / N, j = e % N;
        T[(size_t)i * ld + j] = i == j ? W[(size_t)i * ld + j] : (W[(size_t)i * ld + j] + W[(size_t)j * ld + i]) * 0.5;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < N * N; e += NT) {
        const int i = e / N, j = e % N;
        W[(size_t)i * ld + j] = (i == j) ? 1.0 : 0.0;
    }
    __syncthreads();

    clk.mark(kPhP);
    // ---- Jacobi eigensolve of P, sort (:212-285, :313-316)
    int sweeps, rots;
    jacobi_eigh<NT, MAXN>(N, T, W, ld, sh, sweeps, rots);
    clk.mark(kPhJacobi);
    for (int i = threadIdx.x; i < N; i += NT) sh.diag[i] = T[(size_t)i * ld + i];
    __syncthreads();
    sort_desc<NT, MAXN>(N, sh);
    // sigma_j = max(eig_j, 0)  (:313-316)
    double* sdst = P.s + itm.s_off;
    for (int i = threadIdx.x; i < N; i += NT) sdst[sh.rank[i]] = fmax(sh.diag[i], 0.0);

    // ---- U_r = Q V (columns permuted to sorted order)  (:317)
    wg_gemm<NT>(N, N, N, View{X, ld, 0}, View{W, ld, 0}, [&](int i, int j, double acc) {
        T[(size_t)i * ld + sh.rank[j]] = acc;
    });
    __syncthreads();
    const double oerr = orth_err<NT>(N, T, ld, sh.red);
    clk.mark(kPhU);
    if (!polar_ok || oerr > 1e-8 * rootn) {
        // repair_left_basis_from_right (:185-208): seed = A V S^-1, then MGS.
        double smax = 0.0;
        for (int i = threadIdx.x; i < N; i += NT) smax = fmax(smax, fmax(sh.diag[i], 0.0));
        smax = block_max<NT>(smax, sh.red);
        const double stol = 1e-12 * fmax(smax, 1.0);
        wg_gemm<NT>(N, N, N, asv, View{W, ld, 0}, [&](int i, int j, double acc) {
            const double sj = fmax(sh.diag[j], 0.0);
            X[(size_t)i * ld + sh.rank[j]] = sj > stol ? acc / sj : 0.0;
        });
        __syncthreads();
        gram_schmidt<NT, MAXN>(N, View{X, ld, 0}, T, ld, sh);
        info.repaired = 1;
    }
    clk.mark(kPhRepair);

    // ---- outputs: Vt (or U for wide), then U
    if (!wide) {
        double* vt = P.vt + itm.vt_off;
        for (int e = threadIdx.x; e < N * N; e += NT) {
            const int r = e / N, i = e % N;
            vt[(size_t)sh.rank[i] * N + r] = W[(size_t)r * ld + i];
        }
    } else {
        double* u = P.u + itm.u_off;  // U = (Vt_t)^T : u[r][rank[i]] = V[r][i]
        for (int e = threadIdx.x; e < N * N; e += NT) {
            const int r = e / N, i = e % N;
            u[(size_t)r * N + sh.rank[i]] = W[(size_t)r * ld + i];
        }
    }
    __syncthreads();
    double* Y = T;
    if (tall) {
        Y = arena;  // rows [0, M): disjoint from T (rows [cap-N, cap))
        for (int e = threadIdx.x; e < M * N; e += NT) {
            const int i = e / N, j = e % N;
            Y[(size_t)i * ld + j] = i < N ? T[(size_t)i * ld + j] : 0.0;
        }
        __syncthreads();
        apply_q<NT, MAXN>(M, N, qs, Y, ld, sh);
    }
    if (!wide) {
        double* u = P.u + itm.u_off;
        for (int e = threadIdx.x; e < M * N; e += NT) {
            const int i = e / N, j = e % N;
            u[(size_t)i * N + j] = Y[(size_t)i * ld + j];
        }
    } else {
        double* vt = P.vt + itm.vt_off;  // vt (N x M) = U_t^T
        for (int e = threadIdx.x; e < M * N; e += NT) {
            const int i = e / N, j = e % N;
            vt[(size_t)j * M + i] = Y[(size_t)i * ld + j];
        }
    }
    rots = (int)block_sum<NT>((double)rots, sh.red);
    clk.mark(kPhOut);
    if (threadIdx.x == 0 && P.info) {
        info.jacobi_sweeps = sweeps;
        info.rotations = rots;
        info.path = itm.mode == kModeMicroFallback ? 2 : 0;
        info.micro_fallback = itm.mode == kModeMicroFallback ? 1 : 0;
        P.info[itm.idx] = info;
    }
    __syncthreads();
}

template <int NT, int MAXN, bool LDS_ARENA>
__global__ __launch_bounds__(NT) void small_svd_kernel(const Item* __restrict__ items, const int* count_ptr,
                                                       int* counter, Pools pools, double* gws,
                                                       int64_t gws_stride) {
    extern __shared__ double dyn_lds[];
    __shared__ SmallShared<NT, MAXN> sh;
    double* arena = LDS_ARENA ? dyn_lds : gws + (size_t)blockIdx.x * gws_stride;
    const int count = *count_ptr;
    for (;;) {
        if (threadIdx.x == 0) sh.item = atomicAdd(counter, 1);
        __syncthreads();
        const int it = sh.item;
        __syncthreads();
        if (it >= count) break;
        solve_item<NT, MAXN>(items[it], pools, arena, sh);
    }
}

__global__ void reset_counters_kernel(int* counters, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) counters[i] = 0;
}

// ---- launch table -------------------------------------------------------

template <int NT, int MAXN, bool LDS>
static hipError_t launch_small(const SmallLaunch& L, hipStream_t st) {
    if (LDS && L.lds_bytes > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&small_svd_kernel<NT, MAXN, LDS>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)L.lds_bytes);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((small_svd_kernel<NT, MAXN, LDS>), dim3(L.grid), dim3(NT), L.lds_bytes, st, L.items,
                       L.count_ptr, L.counter, L.pools, L.gws, L.gws_stride);
    return hipGetLastError();
}

template <int NT, int MAXN, bool LDS>
static int max_blocks_small(size_t lds) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, small_svd_kernel<NT, MAXN, LDS>, NT, lds) != hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

#define LSVD_SMALL_VARIANTS(X) \
    X(0, 64, 16, true)         \
    X(1, 128, 32, true)        \
    X(2, 256, 64, true)        \
    X(3, 256, 128, false)      \
    X(4, 256, 1024, false)     \
    X(5, 64, 32, true)         \
    X(6, 128, 64, true)        \
    X(7, 512, 128, false)

const SmallVariant* small_variants(int* count) {
    static const SmallVariant table[] = {
#define X(id, nt, maxn, lds) {id, nt, maxn, lds},
        LSVD_SMALL_VARIANTS(X)
#undef X
    };
    *count = (int)(sizeof(table) / sizeof(table[0]));
    return table;
}

hipError_t launch_small_variant(int id, const SmallLaunch& L, hipStream_t st) {
    if (id >= kOsMulti) return launch_onesided_multi(L, st);
    if (id >= kOsBase) return launch_onesided(id - kOsBase, L, st);
    if (id >= kV2Wide) return launch_small2(id - kV2Wide, true, L, st);
    if (id >= kV2Multi) return launch_small2_multi(L, st);
    if (id >= kV2Base) return launch_small2(id - kV2Base, false, L, st);
    switch (id) {
#define X(vid, nt, maxn, lds) \
    case vid:                 \
        return launch_small<nt, maxn, lds>(L, st);
        LSVD_SMALL_VARIANTS(X)
#undef X
    }
    return hipErrorInvalidValue;
}

int small_variant_blocks_per_cu(int id, size_t lds) {
    if (id >= kOsMulti) return onesided_multi_blocks_per_cu(lds);
    if (id >= kOsBase) return onesided_blocks_per_cu(id - kOsBase, lds);
    if (id >= kV2Wide) return small2_blocks_per_cu(id - kV2Wide, true, lds);
    if (id >= kV2Multi) return small2_multi_blocks_per_cu(lds);
    if (id >= kV2Base) return small2_blocks_per_cu(id - kV2Base, false, lds);
    switch (id) {
#define X(vid, nt, maxn, lds_) \
    case vid:                  \
        return max_blocks_small<nt, maxn, lds_>(lds);
        LSVD_SMALL_VARIANTS(X)
#undef X
    }
    return 1;
}

hipError_t launch_reset_counters(int* counters, int n, hipStream_t st) {
    hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(64), 0, st, counters, n);
    return hipGetLastError();
}

}  // namespace lsvd
