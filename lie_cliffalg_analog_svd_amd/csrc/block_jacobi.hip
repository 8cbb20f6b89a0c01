// block_jacobi.hip -- eigh_jacobi_full (src/lie_svd_small.rs:212-285) of one
// large symmetric matrix across the whole chip, as a block Jacobi method.
//
// The index range 0..n_pad-1 is cut into nb = 2B blocks of b rows; the
// round-robin schedule over blocks pairs them into B disjoint block pairs
// per layer (nb - 1 layers per sweep).  A block pair (I, J) is the index set
// I u J of BB = 2b indices.  Per layer:
//   inner  (one workgroup per block pair): the BB x BB principal submatrix S
//          in LDS takes Jacobi rotations -- the reference's rotation
//          (Rutishauser's tangent, the 1e-300 skip, a_pp - t a_pq /
//          a_qq + t a_pq, zeroed a_pq) applied to all disjoint pairs of an
//          inner layer at once -- and the accumulated rotation W (BB x BB,
//          orthogonal) is stored.  In the first layer of a sweep every pair
//          of I u J is rotated (BB - 1 round-robin inner layers, covering the
//          pairs inside each block); in the other layers only the pairs
//          across I x J (b inner layers, pair (i, b + (i + l) mod b)), since
//          the pairs inside a block were taken in the first layer;
//   update (one launch, one workgroup per tile): A <- G^T A G and V <- V G,
//          where G is
//          W of every block pair on its index set.  Tile (P, Q) of A (rows
//          I u J of pair P, columns of pair Q) becomes W_P^T A_PQ W_Q, two
//          BB^3 products on v_mfma_f64_16x16x4f64 with the operands in LDS;
//          only tiles with P <= Q are computed and each is written to both
//          A_PQ and A_QP (the diagonal tiles from their upper triangle), so A
//          stays exactly symmetric.  V tiles (BB rows x pair P) take V W_P.
// Every global index pair (p, q) is rotated once per sweep, as in the
// scalar cyclic sweep; the order of rotations differs, so results agree
// with the scalar method to rounding.
#include <hip/hip_runtime.h>
#include <math.h>

#include "lsvd_device.hpp"
#include "lsvd_large.hpp"

namespace lsvd {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

// Global index of local index r of block pair `slot` in `layer`.
__device__ __forceinline__ int bj_index(int nb, int b, int layer, int slot, int r) {
    int bi, bj;
    rr_pair(nb, layer, slot, bi, bj);  // nb is even: no phantom block
    return r < b ? bi * b + r : bj * b + (r - b);
}

template <int BB>
constexpr int bj_inner_threads() { return 16 * BB; }

template <int BB>
__global__ __launch_bounds__(16 * BB) void bj_inner_kernel(const double* A, int ld, int nb, int layer, double* Wbuf,
                                                           int* rotations) {
    const bool full = layer == 0;  // pairs inside the blocks as well
    constexpr int b = BB / 2, L = BB + 1, NP = BB / 2, NT = bj_inner_threads<BB>();
    extern __shared__ double lds[];
    double* S = lds;          // BB x L
    double* W = S + BB * L;   // BB x L
    __shared__ double rc[NP], rs[NP], rdp[NP], rdq[NP];
    __shared__ int rp[NP], rq[NP], rf[NP], nrot;
    const int slot = blockIdx.x, tid = threadIdx.x;
    if (tid == 0) nrot = 0;
    constexpr int PER = BB * BB / NT;
    static_assert(BB * BB % NT == 0, "tile entries per thread");
    double ra[PER];  // all loads in flight before the first LDS store
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int e = tid + i * NT, r = e / BB, c = e % BB;
        const int gr = bj_index(nb, b, layer, slot, r), gc = bj_index(nb, b, layer, slot, c);
        ra[i] = A[(size_t)gr * ld + gc];
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int e = tid + i * NT, r = e / BB, c = e % BB;
        S[r * L + c] = ra[i];
        W[r * L + c] = r == c ? 1.0 : 0.0;
    }
    __syncthreads();
    int mine = 0;
    const int inner_layers = full ? BB - 1 : b;
    for (int il = 0; il < inner_layers; ++il) {
        if (tid < NP) {
            int p, q;
            if (full) {
                rr_pair(BB, il, tid, p, q);
            } else {
                p = tid;
                q = b + (tid + il) % b;
            }
            const double apq = S[p * L + q], app = S[p * L + p], aqq = S[q * L + q];
            double c, s, t;
            const int f = jacobi_angle(app, aqq, apq, true, c, s, t);  // (:229-235)
            if (f) {
                rdp[tid] = app - t * apq;  // (:236-241)
                rdq[tid] = aqq + t * apq;
                ++mine;
            }
            rc[tid] = c;
            rs[tid] = s;
            rp[tid] = p;
            rq[tid] = q;
            rf[tid] = f;
        }
        __syncthreads();
        // S <- J^T S J on every 2 x 2 block (rows of pair k1, columns of pair
        // k2) in one pass: the row rotation of k1, then the column rotation
        // of k2 (:242-257); the rotated pair's own block takes the
        // reference's exact values (a_pq = a_qp = 0).  W <- W J on columns.
        for (int e = tid; e < NP * NP; e += NT) {
            const int k1 = e / NP, k2 = e % NP;
            const int f1 = rf[k1], f2 = rf[k2];
            const int p1 = rp[k1], q1 = rq[k1];
            if (k1 == k2) {
                if (f1) {
                    S[p1 * L + p1] = rdp[k1];
                    S[p1 * L + q1] = 0.0;
                    S[q1 * L + p1] = 0.0;
                    S[q1 * L + q1] = rdq[k1];
                }
                continue;
            }
            if (!f1 && !f2) continue;
            const int p2 = rp[k2], q2 = rq[k2];
            const double c1 = rc[k1], s1 = rs[k1], c2 = rc[k2], s2 = rs[k2];
            double* xp = S + p1 * L;
            double* xq = S + q1 * L;
            const double xpp = xp[p2], xpq = xp[q2], xqp = xq[p2], xqq = xq[q2];
            const double rpp = c1 * xpp - s1 * xqp, rpq = c1 * xpq - s1 * xqq;
            const double rqp = s1 * xpp + c1 * xqp, rqq = s1 * xpq + c1 * xqq;
            xp[p2] = c2 * rpp - s2 * rpq;
            xp[q2] = s2 * rpp + c2 * rpq;
            xq[p2] = c2 * rqp - s2 * rqq;
            xq[q2] = s2 * rqp + c2 * rqq;
        }
        for (int e = tid; e < NP * BB; e += NT) {
            const int k = e / BB, row = e % BB;
            if (!rf[k]) continue;
            const double c = rc[k], s = rs[k];
            double* wp = W + row * L + rp[k];
            double* wq = W + row * L + rq[k];
            const double up = *wp, uq = *wq;
            *wp = c * up - s * uq;
            *wq = s * up + c * uq;
        }
        __syncthreads();
    }
    for (int e = tid; e < BB * BB; e += NT) {
        const int r = e / BB, c = e % BB;
        Wbuf[((size_t)slot * BB + r) * BB + c] = W[r * L + c];
    }
    if (mine) atomicAdd(&nrot, mine);
    __syncthreads();
    if (tid == 0 && nrot) atomicAdd(rotations, nrot);
}

// C (BB x BB, LDS, ldc) = op(X) Y with X, Y in LDS (ld L); op(X) = X^T when
// TX.  Wave w takes output tiles w, w + 4, ... of the (BB/16)^2 tiles.
template <int BB, bool TX>
__device__ __forceinline__ void bj_tile_gemm(const double* X, const double* Y, double* Cm, int L) {
    constexpr int NT = BB / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = lane & 15, kk = lane >> 4;
    for (int t = wave; t < NT * NT; t += 4) {
        const int I = t / NT, J = t % NT;
        d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
        for (int k0 = 0; k0 < BB; k0 += 4) {
            const int k = k0 + kk;
            const double xv = TX ? X[k * L + I * 16 + i] : X[(I * 16 + i) * L + k];
            const double yv = Y[k * L + J * 16 + i];
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(xv, yv, acc, 0, 0, 0);
        }
        // accumulator element q of a lane: row 4 q + lane / 16, column lane % 16
#pragma unroll
        for (int q = 0; q < 4; ++q) Cm[(I * 16 + 4 * q + kk) * L + J * 16 + i] = acc[q];
    }
}

// One launch per layer: workgroups 0 .. B (B + 1) / 2 - 1 take the A tiles
// (P <= Q), the next (n_pad / BB) B the V tiles (rows R, pair P's columns).
template <int BB>
__global__ __launch_bounds__(256) void bj_update_kernel(double* A, double* V, int ld, int nb, int layer,
                                                        const double* Wbuf) {
    constexpr int b = BB / 2, L = BB + 1, UT = 256, PER = BB * BB / UT;
    static_assert(BB * BB % UT == 0, "tile entries per thread");
    extern __shared__ double lds[];
    double* X = lds;            // tile of A or V
    double* WP = X + BB * L;
    double* T = WP + BB * L;
    double* WQ = T + BB * L;    // A tiles only
    const int B = nb / 2;
    const int nA = B * (B + 1) / 2;
    const int tid = threadIdx.x;
    int t = blockIdx.x;
    if (t < nA) {
        // upper-triangle tile t -> (P, Q), P <= Q
        int P = 0;
        while (t >= B - P) {
            t -= B - P;
            ++P;
        }
        const int Q = P + t;
        // all loads issued before the first LDS store, so their latencies overlap
        double ra[PER], rp[PER], rq[PER];
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int e = tid + i * UT, r = e / BB, c = e % BB;
            const int gr = bj_index(nb, b, layer, P, r), gc = bj_index(nb, b, layer, Q, c);
            ra[i] = A[(size_t)gr * ld + gc];
            rp[i] = Wbuf[((size_t)P * BB + r) * BB + c];
            rq[i] = Wbuf[((size_t)Q * BB + r) * BB + c];
        }
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int e = tid + i * UT, r = e / BB, c = e % BB;
            X[r * L + c] = ra[i];
            WP[r * L + c] = rp[i];
            WQ[r * L + c] = rq[i];
        }
        __syncthreads();
        bj_tile_gemm<BB, false>(X, WQ, T, L);  // T = A_PQ W_Q
        __syncthreads();
        bj_tile_gemm<BB, true>(WP, T, X, L);   // X = W_P^T T
        __syncthreads();
        for (int e = tid; e < BB * BB; e += UT) {
            const int r = e / BB, c = e % BB;
            const int gr = bj_index(nb, b, layer, P, r), gc = bj_index(nb, b, layer, Q, c);
            double v;
            if (P == Q) v = r <= c ? X[r * L + c] : X[c * L + r];
            else v = X[r * L + c];
            A[(size_t)gr * ld + gc] = v;
            if (P != Q) {
                // A_QP = (A_PQ)^T: row gc' = pair Q's r-th index, column pair P's c-th
                const int hr = bj_index(nb, b, layer, Q, r), hc = bj_index(nb, b, layer, P, c);
                A[(size_t)hr * ld + hc] = X[c * L + r];
            }
        }
        return;
    }
    t -= nA;
    const int P = t % B, R = t / B;  // V rows R*BB .. R*BB + BB - 1, pair P's columns
    if (R >= ld / BB) return;
    double rv[PER], rp[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int e = tid + i * UT, r = e / BB, c = e % BB;
        const int gc = bj_index(nb, b, layer, P, c);
        rv[i] = V[(size_t)(R * BB + r) * ld + gc];
        rp[i] = Wbuf[((size_t)P * BB + r) * BB + c];
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int e = tid + i * UT, r = e / BB, c = e % BB;
        X[r * L + c] = rv[i];
        WP[r * L + c] = rp[i];
    }
    __syncthreads();
    bj_tile_gemm<BB, false>(X, WP, T, L);  // T = V_RP W_P
    __syncthreads();
    for (int e = tid; e < BB * BB; e += UT) {
        const int r = e / BB, c = e % BB;
        const int gc = bj_index(nb, b, layer, P, c);
        V[(size_t)(R * BB + r) * ld + gc] = T[r * L + c];
    }
}

// A_pad = [P 0; 0 0] (n_pad x n_pad), V_pad = I.
__global__ void bj_setup_kernel(const double* P, int n, int ld, double* A, double* V) {
    const size_t tot = (size_t)ld * ld;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(e / ld), c = (int)(e % ld);
        A[e] = (r < n && c < n) ? P[(size_t)r * n + c] : 0.0;
        V[e] = r == c ? 1.0 : 0.0;
    }
}

// Upper off-diagonal sum of squares in two fixed-shape passes (per-block
// partials, then one block), so every run sees the same bits.
__global__ void bj_off_kernel(const double* A, int n, int ld, double* partials) {
    __shared__ double red[4];
    double s = 0.0;
    for (int r = blockIdx.x; r < n; r += gridDim.x)
        for (int c = r + 1 + threadIdx.x; c < n; c += blockDim.x) {
            const double v = A[(size_t)r * ld + c];
            s += v * v;
        }
    s = block_sum<256>(s, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ void bj_off_final_kernel(const double* partials, int np, double* out) {
    __shared__ double red[4];
    double s = threadIdx.x < np ? partials[threadIdx.x] : 0.0;
    s = block_sum<256>(s, red);
    if (threadIdx.x == 0) *out = s;
}

// d[i] = A[i][i]; Vc (n x n) = V[0:n, 0:n].
__global__ void bj_extract_kernel(const double* A, const double* V, int n, int ld, double* d, double* Vc) {
    const size_t tot = (size_t)n * n;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(e / n), c = (int)(e % n);
        Vc[e] = V[(size_t)r * ld + c];
        if (r == c) d[r] = A[(size_t)r * ld + r];
    }
}

template <int BB>
constexpr size_t bj_inner_lds() { return sizeof(double) * 2 * BB * (BB + 1); }
template <int BB>
constexpr size_t bj_update_lds() { return sizeof(double) * 4 * BB * (BB + 1); }

template <int BB>
hipError_t bj_prepare_bb() {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&bj_inner_kernel<BB>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)bj_inner_lds<BB>());
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute(reinterpret_cast<const void*>(&bj_update_kernel<BB>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)bj_update_lds<BB>());
}

template <int BB>
hipError_t bj_launch_bb(const BlockJacobi& j, bool inner, int layer, double* W, hipStream_t st) {
    const int B = j.nb / 2;
    if (inner) {
        hipLaunchKernelGGL(bj_inner_kernel<BB>, dim3(B), dim3(bj_inner_threads<BB>()), bj_inner_lds<BB>(), st, j.A,
                           j.ld, j.nb, layer, W, j.rotations);
    } else {
        hipLaunchKernelGGL(bj_update_kernel<BB>, dim3(B * (B + 1) / 2 + (j.ld / BB) * B), dim3(256),
                           bj_update_lds<BB>(), st, j.A, j.V, j.ld, j.nb, layer, W);
    }
    return hipGetLastError();
}

hipError_t bj_launch(const BlockJacobi& j, bool inner, int layer, double* W, hipStream_t st) {
    switch (j.bb) {
        case 16: return bj_launch_bb<16>(j, inner, layer, W, st);
        case 32: return bj_launch_bb<32>(j, inner, layer, W, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

int bj_tile_size(int n) { return n <= 256 ? 16 : 32; }

hipError_t bj_setup(const BlockJacobi& j, const double* P, hipStream_t st) {
    if (j.bb != 16 && j.bb != 32) return hipErrorInvalidValue;
    hipError_t e = j.bb == 16 ? bj_prepare_bb<16>() : bj_prepare_bb<32>();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(bj_setup_kernel, dim3(512), dim3(256), 0, st, P, j.n, j.ld, j.A, j.V);
    return hipGetLastError();
}

hipError_t bj_inner(const BlockJacobi& j, int layer, double* W, hipStream_t st) {
    return bj_launch(j, true, layer, W, st);
}
hipError_t bj_update(const BlockJacobi& j, int layer, const double* W, hipStream_t st) {
    return bj_launch(j, false, layer, const_cast<double*>(W), st);
}

hipError_t bj_offnorm(const BlockJacobi& j, double* out, hipStream_t st) {
    hipLaunchKernelGGL(bj_off_kernel, dim3(kBjOffBlocks), dim3(256), 0, st, j.A, j.n, j.ld, j.partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(bj_off_final_kernel, dim3(1), dim3(256), 0, st, j.partials, kBjOffBlocks, out);
    return hipGetLastError();
}

hipError_t bj_extract(const BlockJacobi& j, double* d, double* Vc, hipStream_t st) {
    hipLaunchKernelGGL(bj_extract_kernel, dim3(512), dim3(256), 0, st, j.A, j.V, j.n, j.ld, d, Vc);
    return hipGetLastError();
}

}  // namespace lsvd
