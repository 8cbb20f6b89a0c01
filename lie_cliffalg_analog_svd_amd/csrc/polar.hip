// polar.hip -- MFMA GEMMs of the Newton-Schulz polar iteration for large
// matrices (newton_schulz_polar, src/lie_svd_small.rs:158-183, generalised
// to tall m x n), plus the small helper kernels of the iteration.
//
// One NS step is two products with fused epilogues:
//   M     = 3I - X^T X      (n x n, symmetric: only upper tiles are computed,
//                            the epilogue writes both halves)
//   X_new = 0.5 X M         (m x n; the epilogue accumulates ||X_new - X||_F^2
//                            and a non-finite flag for the stopping test)
// Three arithmetic variants share the tiling:
//   f64  : v_mfma_f64_16x16x4f64, FP64 in LDS           (plain path, refinement)
//   f32  : v_mfma_f32_16x16x4f32, FP32 in LDS
//   bf16 : v_mfma_f32_16x16x32_bf16, FP32 data rounded to BF16 when staged
//          into LDS, FP32 accumulation
// Workgroup: 256 threads = 2 x 2 wavefronts; each wavefront owns a
// (BM/2) x (BN/2) block of 16 x 16 MFMA tiles.
#include <hip/hip_runtime.h>
#include <math.h>

#include "lsvd_device.hpp"
#include "lsvd_polar.hpp"

namespace lsvd {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

template <class T>
struct Acc4;
template <>
struct Acc4<double> {
    typedef d4 type;
};
template <>
struct Acc4<float> {
    typedef f4 type;
};

// Row of accumulator register r in lane `lane` within a 16 x 16 MFMA tile
// (column = lane % 16).  Measured on gfx950: the FP64 16x16x4 instruction
// interleaves rows across lane groups (row = 4 r + lane / 16), the FP32 and
// BF16 16x16 instructions keep 4 consecutive rows per lane group
// (row = 4 (lane / 16) + r).
template <class T>
__device__ __forceinline__ int acc_row(int lane, int r);
template <>
__device__ __forceinline__ int acc_row<double>(int lane, int r) { return 4 * r + (lane >> 4); }
template <>
__device__ __forceinline__ int acc_row<float>(int lane, int r) { return 4 * (lane >> 4) + r; }

__device__ __forceinline__ d4 mfma4(double a, double b, d4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
__device__ __forceinline__ f4 mfma4(float a, float b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

// Epilogue: element (i, j) of the product, value `acc` (already in T).
struct EpiArgs {
    void* c;          // output (T*), ldc = N
    const void* xold; // kHalfDelta: previous iterate (T*), same shape as c
    double* sum;      // kHalfDelta / kOrth: accumulated squared norm
    int* bad;         // kHalfDelta: non-finite flag
    double alpha;
};

template <class T, int EPI>
__device__ __forceinline__ void epi_apply(const EpiArgs& e, int N, int i, int j, T acc, double& part, int& bad,
                                          bool mirror) {
    T* c = reinterpret_cast<T*>(e.c);
    if (EPI == kEpiStore) {
        const T v = (T)(e.alpha * (double)acc);
        c[(size_t)i * N + j] = v;
        if (mirror) c[(size_t)j * N + i] = v;
    } else if (EPI == kEpi3IMinus) {
        const T v = (T)((i == j ? e.alpha : 0.0) - (double)acc);
        c[(size_t)i * N + j] = v;
        if (mirror) c[(size_t)j * N + i] = v;
        if (e.sum) {  // ||X^T X - I||_F^2 as a by-product
            const double d = (double)acc - (i == j ? 1.0 : 0.0);
            part += (mirror ? 2.0 : 1.0) * d * d;
        }
    } else if (EPI == kEpiHalfDelta) {
        const T v = (T)(e.alpha * (double)acc);
        const T* xo = reinterpret_cast<const T*>(e.xold);
        const double d = (double)v - (double)xo[(size_t)i * N + j];
        c[(size_t)i * N + j] = v;
        part += d * d;
        bad |= !isfinite((double)v);
    } else {  // kEpiOrth: ||X^T X - I||_F^2, symmetric tiles count twice off the diagonal
        const double d = (double)acc - (i == j ? 1.0 : 0.0);
        part += (mirror ? 2.0 : 1.0) * d * d;
    }
}

template <int EPI>
__device__ __forceinline__ void epi_finish(const EpiArgs& e, double part, int bad) {
    if (EPI != kEpiHalfDelta && EPI != kEpiOrth && EPI != kEpi3IMinus) return;
    if (EPI == kEpi3IMinus && !e.sum) return;  // uniform: kernel argument
    __shared__ double red[4];
    part = block_sum<256>(part, red);
    if (EPI == kEpiHalfDelta) bad = block_or<256>(bad, red);
    if (threadIdx.x == 0) {
        atomicAdd(e.sum, part);
        if (EPI == kEpiHalfDelta && bad) atomicOr(e.bad, 1);
    }
}

// Maps a linear block id to an upper-triangular tile pair (ti <= tj).
__device__ __forceinline__ void tri_tile(int b, int& ti, int& tj) {
    int j = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
    while (j * (j + 1) / 2 > b) --j;
    while ((j + 1) * (j + 2) / 2 <= b) ++j;
    tj = j;
    ti = b - j * (j + 1) / 2;
}

// ---- f64 / f32: C (M x N) = op(A) B, A stored K x M (TRANS_A) or M x K --------
// BM = BN = TILE (64, or 32 for products too small to fill the chip with
// 64-wide tiles), BK = 1024 / TILE, so every thread loads four entries of
// each operand per K tile; wave w owns the (TILE/2)^2 quadrant (w / 2, w % 2).
// SYM: square symmetric output, upper tiles only.
template <class T, bool TRANS_A, bool SYM, int EPI, int TILE>
__global__ __launch_bounds__(256) void gemm_mfma4_kernel(int M, int N, int K, const T* __restrict__ A,
                                                          const T* __restrict__ B, EpiArgs e) {
    constexpr int BM = TILE, BN = TILE, BK = 1024 / TILE, LDA = BM + 1, LDB = BN + 1;
    constexpr int HM = TILE / 2, NS = TILE / 32;  // quadrant edge, 16 x 16 blocks per quadrant edge
    static_assert(TILE == 32 || TILE == 64, "tile edge");
    __shared__ T As[BK * LDA];
    __shared__ T Bs[BK * LDB];
    typedef typename Acc4<T>::type acc_t;
    int ti, tj;
    if (SYM) {
        tri_tile(blockIdx.x, ti, tj);
    } else {
        ti = blockIdx.y;
        tj = blockIdx.x;
    }
    const int i0 = ti * BM, j0 = tj * BN;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w >> 1, wn = w & 1;
    acc_t acc[NS][NS];
#pragma unroll
    for (int a = 0; a < NS; ++a)
#pragma unroll
        for (int b = 0; b < NS; ++b) acc[a][b] = acc_t{0, 0, 0, 0};
    // The next K tile is read from global memory into registers while the
    // current one is multiplied out of LDS.
    T ra[4], rb[4];
    auto gload = [&](int k0) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int e_ = t + 256 * r;
            if (TRANS_A) {  // op(A)[i][k] = A[k][i], rows of A contiguous in i
                const int kk = e_ / BM, ii = e_ % BM;
                const int gk = k0 + kk, gi = i0 + ii;
                ra[r] = (gk < K && gi < M) ? A[(size_t)gk * M + gi] : (T)0;
            } else {
                const int ii = e_ / BK, kk = e_ % BK;
                const int gk = k0 + kk, gi = i0 + ii;
                ra[r] = (gk < K && gi < M) ? A[(size_t)gi * K + gk] : (T)0;
            }
            const int kk = e_ / BN, jj = e_ % BN;
            const int gk = k0 + kk, gj = j0 + jj;
            rb[r] = (gk < K && gj < N) ? B[(size_t)gk * N + gj] : (T)0;
        }
    };
    gload(0);
    for (int k0 = 0; k0 < K; k0 += BK) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int e_ = t + 256 * r;
            if (TRANS_A) {
                As[(e_ / BM) * LDA + e_ % BM] = ra[r];
            } else {
                As[(e_ % BK) * LDA + e_ / BK] = ra[r];
            }
            Bs[(e_ / BN) * LDB + e_ % BN] = rb[r];
        }
        __syncthreads();
        if (k0 + BK < K) gload(k0 + BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            const int kr = kk + (lane >> 4), c16 = lane & 15;
            T av[NS], bv[NS];
#pragma unroll
            for (int a = 0; a < NS; ++a) av[a] = As[kr * LDA + wm * HM + a * 16 + c16];
#pragma unroll
            for (int b = 0; b < NS; ++b) bv[b] = Bs[kr * LDB + wn * HM + b * 16 + c16];
#pragma unroll
            for (int a = 0; a < NS; ++a)
#pragma unroll
                for (int b = 0; b < NS; ++b) acc[a][b] = mfma4(av[a], bv[b], acc[a][b]);
        }
        __syncthreads();
    }
    double part = 0.0;
    int bad = 0;
#pragma unroll
    for (int a = 0; a < NS; ++a)
#pragma unroll
        for (int b = 0; b < NS; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + wm * HM + a * 16 + acc_row<T>(lane, r);
                const int j = j0 + wn * HM + b * 16 + (lane & 15);
                if (i < M && j < N && (!SYM || ti < tj || i <= j))
                    epi_apply<T, EPI>(e, N, i, j, acc[a][b][r], part, bad, SYM && i != j);
            }
    epi_finish<EPI>(e, part, bad);
}

// ---- bf16 operands, f32 data and accumulation --------------------------------
// BM = BN = 128, BK = 32; LDS tiles are [row][k] with k contiguous so one lane
// reads its 8 consecutive k values as one 16-byte load.
template <bool TRANS_A, bool SYM, int EPI>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(int M, int N, int K, const float* __restrict__ A,
                                                         const float* __restrict__ B, EpiArgs e) {
    constexpr int BM = 128, BN = 128, BK = 32, LDK = BK + 8;
    __shared__ __align__(16) __bf16 As[BM * LDK];
    __shared__ __align__(16) __bf16 Bs[BN * LDK];
    int ti, tj;
    if (SYM) {
        tri_tile(blockIdx.x, ti, tj);
    } else {
        ti = blockIdx.y;
        tj = blockIdx.x;
    }
    const int i0 = ti * BM, j0 = tj * BN;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w >> 1, wn = w & 1;
    f4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = f4{0, 0, 0, 0};
    // Each thread owns two (row, 8-wide k chunk) items of each operand tile:
    // it reads the chunk from global memory (float4 pairs where k is
    // contiguous, otherwise one float per k with the lanes of a wave on
    // consecutive rows), and stores it to LDS as one packed 16-byte write.
    // The next K tile is read into registers while the current one is
    // multiplied out of LDS.
    float ra[2][8], rb[2][8];
    auto gload = [&](int k0) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int id = t + 256 * h;
            if (TRANS_A) {  // A stored K x M: i contiguous
                const int ii = id % BM, c = id / BM, gi = i0 + ii;
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const int gk = k0 + 8 * c + q;
                    ra[h][q] = (gk < K && gi < M) ? A[(size_t)gk * M + gi] : 0.0f;
                }
            } else {  // A stored M x K: k contiguous
                const int ii = id / 4, c = id % 4, gi = i0 + ii, gk = k0 + 8 * c;
                if (gi < M && gk + 8 <= K && (K & 3) == 0) {
                    const float4 v0 = *reinterpret_cast<const float4*>(&A[(size_t)gi * K + gk]);
                    const float4 v1 = *reinterpret_cast<const float4*>(&A[(size_t)gi * K + gk + 4]);
                    ra[h][0] = v0.x; ra[h][1] = v0.y; ra[h][2] = v0.z; ra[h][3] = v0.w;
                    ra[h][4] = v1.x; ra[h][5] = v1.y; ra[h][6] = v1.z; ra[h][7] = v1.w;
                } else {
#pragma unroll
                    for (int q = 0; q < 8; ++q) ra[h][q] = (gi < M && gk + q < K) ? A[(size_t)gi * K + gk + q] : 0.0f;
                }
            }
            const int jj = id % BN, c = id / BN, gj = j0 + jj;  // B stored K x N: j contiguous
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int gk = k0 + 8 * c + q;
                rb[h][q] = (gk < K && gj < N) ? B[(size_t)gk * N + gj] : 0.0f;
            }
        }
    };
    gload(0);
    for (int k0 = 0; k0 < K; k0 += BK) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int id = t + 256 * h;
            const int ar = TRANS_A ? id % BM : id / 4, ac = TRANS_A ? id / BM : id % 4;
            bf8 pa, pb;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                pa[q] = (__bf16)ra[h][q];
                pb[q] = (__bf16)rb[h][q];
            }
            *reinterpret_cast<bf8*>(&As[ar * LDK + 8 * ac]) = pa;
            *reinterpret_cast<bf8*>(&Bs[(id % BN) * LDK + 8 * (id / BN)]) = pb;
        }
        __syncthreads();
        if (k0 + BK < K) gload(k0 + BK);
        const int kq = 8 * (lane >> 4), c16 = lane & 15;
        bf8 av[4], bv[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) av[a] = *reinterpret_cast<const bf8*>(&As[(wm * 64 + a * 16 + c16) * LDK + kq]);
#pragma unroll
        for (int b = 0; b < 4; ++b) bv[b] = *reinterpret_cast<const bf8*>(&Bs[(wn * 64 + b * 16 + c16) * LDK + kq]);
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a], bv[b], acc[a][b], 0, 0, 0);
        __syncthreads();
    }
    double part = 0.0;
    int bad = 0;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + wm * 64 + a * 16 + acc_row<float>(lane, r);
                const int j = j0 + wn * 64 + b * 16 + (lane & 15);
                if (i < M && j < N && (!SYM || ti < tj || i <= j))
                    epi_apply<float, EPI>(e, N, i, j, acc[a][b][r], part, bad, SYM && i != j);
            }
    epi_finish<EPI>(e, part, bad);
}

// ---- helpers --------------------------------------------------------------------
// ||x||^2 in two fixed-shape passes (per-block partials, then one block
// sums them in a fixed order), so the norm that scales X0 is the same bits
// on every run.
__global__ void sumsq_kernel(const double* x, size_t n, double* partials) {
    __shared__ double red[4];
    double s = 0.0;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        s += x[e] * x[e];
    s = block_sum<256>(s, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ void sumsq_final_kernel(const double* partials, double* out) {
    __shared__ double red[4];
    double s = partials[threadIdx.x];
    s = block_sum<256>(s, red);
    if (threadIdx.x == 0) *out = s;
}

// ||B - B^T||_F^2 of an n x n matrix: every ordered pair (i, j) once, so
// each skew entry is counted twice as in the Frobenius norm; a 32 x 32
// tile and its mirror are staged in LDS so both reads coalesce.
__global__ __launch_bounds__(256) void skew_sumsq_kernel(const double* b, int n, double* partials) {
    __shared__ double t[32][33];
    __shared__ double red[4];
    const int tiles = (n + 31) / 32;
    double s = 0.0;
    for (int tIdx = blockIdx.x; tIdx < tiles * tiles; tIdx += gridDim.x) {
        const int I = tIdx / tiles, J = tIdx % tiles;
        __syncthreads();
        for (int e = threadIdx.x; e < 32 * 32; e += 256) {  // mirror tile (J, I), stored transposed
            const int r = e / 32, c = e % 32, gi = J * 32 + r, gj = I * 32 + c;
            t[c][r] = (gi < n && gj < n) ? b[(size_t)gi * n + gj] : 0.0;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < 32 * 32; e += 256) {
            const int r = e / 32, c = e % 32, gi = I * 32 + r, gj = J * 32 + c;
            if (gi < n && gj < n) {
                const double d = b[(size_t)gi * n + gj] - t[r][c];
                s += d * d;
            }
        }
    }
    s = block_sum<256>(s, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// y = x / max(sqrt(*sumsq), 1e-300), into f64 or f32 storage
template <class T>
__global__ void scale_kernel(const double* x, size_t n, const double* sumsq, T* y) {
    const double s = fmax(sqrt(*sumsq), 1e-300);
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        y[e] = (T)(x[e] / s);
}

__global__ void f32_to_f64_kernel(const float* x, size_t n, double* y) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        y[e] = (double)x[e];
}

__global__ void zero_kernel(double* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0.0;
}

// ---- launch wrappers ---------------------------------------------------------------
template <int TILE>
static dim3 grid4(int M, int N, bool sym) {
    if (sym) {
        const int tn = (N + TILE - 1) / TILE;
        return dim3(tn * (tn + 1) / 2);
    }
    return dim3((N + TILE - 1) / TILE, (M + TILE - 1) / TILE);
}

// 32-wide tiles when 64-wide ones would give fewer than four workgroups per
// CU (256 CUs; squares up to ~2048): four times the workgroups at half the
// reuse per LDS load.  Measured on the plain polar iteration: 1024^2 4.9 ->
// 4.2 ms, 1536^2 13.4 -> 10.0 ms against a threshold of one per CU; two per
// CU and eight per CU were both slower.
constexpr int kGemmFillWgs = 1024;

template <class T, bool TA, bool SYM, int EPI>
static hipError_t launch4(int M, int N, int K, const T* A, const T* B, const EpiArgs& e, hipStream_t st) {
    const dim3 g64 = grid4<64>(M, N, SYM);
    if ((int)(g64.x * g64.y) < kGemmFillWgs) {
        hipLaunchKernelGGL((gemm_mfma4_kernel<T, TA, SYM, EPI, 32>), grid4<32>(M, N, SYM), dim3(256), 0, st, M, N, K,
                           A, B, e);
    } else {
        hipLaunchKernelGGL((gemm_mfma4_kernel<T, TA, SYM, EPI, 64>), g64, dim3(256), 0, st, M, N, K, A, B, e);
    }
    return hipGetLastError();
}

template <bool TA, bool SYM, int EPI>
static hipError_t launch_bf16(int M, int N, int K, const float* A, const float* B, const EpiArgs& e, hipStream_t st) {
    dim3 grid;
    if (SYM) {
        const int tn = (N + 127) / 128;
        grid = dim3(tn * (tn + 1) / 2);
    } else {
        grid = dim3((N + 127) / 128, (M + 127) / 128);
    }
    hipLaunchKernelGGL((gemm_bf16_kernel<TA, SYM, EPI>), grid, dim3(256), 0, st, M, N, K, A, B, e);
    return hipGetLastError();
}

// M = 3I - X^T X (n x n) from X (m x n); *orth_sq += ||X^T X - I||_F^2 when non-null
hipError_t ns_gram(int prec, int m, int n, const void* X, void* Mout, double* orth_sq, hipStream_t st,
                   double alpha) {
    EpiArgs e{Mout, nullptr, orth_sq, nullptr, alpha};
    if (prec == kPolF64) return launch4<double, true, true, kEpi3IMinus>(n, n, m, (const double*)X, (const double*)X, e, st);
    if (prec == kPolF32) return launch4<float, true, true, kEpi3IMinus>(n, n, m, (const float*)X, (const float*)X, e, st);
    return launch_bf16<true, true, kEpi3IMinus>(n, n, m, (const float*)X, (const float*)X, e, st);
}

// Xn = c X M, accumulating ||Xn - X||^2 into *dsq and non-finite into *bad
hipError_t ns_update(int prec, int m, int n, const void* X, const void* Mm, void* Xn, double* dsq, int* bad,
                     hipStream_t st, double c) {
    EpiArgs e{Xn, X, dsq, bad, c};
    if (prec == kPolF64) return launch4<double, false, false, kEpiHalfDelta>(m, n, n, (const double*)X, (const double*)Mm, e, st);
    if (prec == kPolF32) return launch4<float, false, false, kEpiHalfDelta>(m, n, n, (const float*)X, (const float*)Mm, e, st);
    return launch_bf16<false, false, kEpiHalfDelta>(m, n, n, (const float*)X, (const float*)Mm, e, st);
}

// *sum += ||X^T X - I||_F^2 (FP64)
hipError_t ns_orth(int m, int n, const double* X, double* sum, hipStream_t st) {
    EpiArgs e{nullptr, nullptr, sum, nullptr, 1.0};
    return launch4<double, true, true, kEpiOrth>(n, n, m, X, X, e, st);
}

// Plain C = op(A) B in FP64 (tests and the adjoint)
hipError_t gemm_f64(bool trans_a, int M, int N, int K, const double* A, const double* B, double* C, double alpha,
                    hipStream_t st) {
    EpiArgs e{C, nullptr, nullptr, nullptr, alpha};
    if (trans_a) return launch4<double, true, false, kEpiStore>(M, N, K, A, B, e, st);
    return launch4<double, false, false, kEpiStore>(M, N, K, A, B, e, st);
}

hipError_t gemm_lowp(int prec, bool trans_a, int M, int N, int K, const float* A, const float* B, float* C,
                     hipStream_t st) {
    EpiArgs e{C, nullptr, nullptr, nullptr, 1.0};
    if (prec == kPolF32) {
        if (trans_a) return launch4<float, true, false, kEpiStore>(M, N, K, A, B, e, st);
        return launch4<float, false, false, kEpiStore>(M, N, K, A, B, e, st);
    }
    if (trans_a) return launch_bf16<true, false, kEpiStore>(M, N, K, A, B, e, st);
    return launch_bf16<false, false, kEpiStore>(M, N, K, A, B, e, st);
}

hipError_t polar_sumsq(const double* x, size_t n, double* out, double* partials, hipStream_t st) {
    hipLaunchKernelGGL(sumsq_kernel, dim3(kSumsqBlocks), dim3(256), 0, st, x, n, partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(sumsq_final_kernel, dim3(1), dim3(kSumsqBlocks), 0, st, partials, out);
    return hipGetLastError();
}

hipError_t polar_skew_sumsq(const double* b, int n, double* out, double* partials, hipStream_t st) {
    hipLaunchKernelGGL(skew_sumsq_kernel, dim3(kSumsqBlocks), dim3(256), 0, st, b, n, partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(sumsq_final_kernel, dim3(1), dim3(kSumsqBlocks), 0, st, partials, out);
    return hipGetLastError();
}

hipError_t polar_scale(const double* x, size_t n, const double* sumsq, void* y, bool f32, hipStream_t st) {
    if (f32)
        hipLaunchKernelGGL(scale_kernel<float>, dim3(1024), dim3(256), 0, st, x, n, sumsq, (float*)y);
    else
        hipLaunchKernelGGL(scale_kernel<double>, dim3(1024), dim3(256), 0, st, x, n, sumsq, (double*)y);
    return hipGetLastError();
}

hipError_t polar_widen(const float* x, size_t n, double* y, hipStream_t st) {
    hipLaunchKernelGGL(f32_to_f64_kernel, dim3(1024), dim3(256), 0, st, x, n, y);
    return hipGetLastError();
}

hipError_t polar_zero(double* p, int n, hipStream_t st) {
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, st, p, n);
    return hipGetLastError();
}

}  // namespace lsvd
