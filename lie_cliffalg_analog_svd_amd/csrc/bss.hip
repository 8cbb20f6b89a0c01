// bss.hip -- the dense steps of LieSvdBss::separate (src/lie_svd_bss.rs:
// 58-101) for a batch of channels x samples observation blocks:
//   channel means and the covariance X_c X_c^T / S      (:60-63, :162-171)
//   the whitening matrix from LieSvdSmall's (U, sigma)   (:66-73)
//   whitened = W X_c and separated = V^T whitened        (:74, :81)
//   lagged covariances of the whitened channels          (:75, :173-197)
//   unmixing = V^T W and the channel coherence           (:82-84, :104-126)
// The covariance and the lagged covariances are one kernel: cross-products
// of a channel block with its lagged copy, staged through LDS in chunks of
// 32 samples at 64 channels (proportionally more at fewer channels) and split
// over the samples into about 8192 workgroups in all, so that the chunks of a
// block are staged side by side rather than one after another; the split sums
// are added in a fixed order by the reduce kernel.
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>

#include "lsvd_bss.hpp"
#include "lsvd_device.hpp"

namespace lsvd {
namespace {

// LDS of one staged operand: 64 channels x 33 (32 samples + pad) doubles;
// fewer channels stage proportionally more samples per step.
constexpr int kStage = kBssMaxChannels * 33;
constexpr int kAccPerThread = kBssMaxChannels * kBssMaxChannels / 256;

// Samples staged per step for `ch` channels: a multiple of 32 with
// ch (chunk + 1) <= kStage.
inline int bss_chunk(int ch) { return std::max(32, (kStage / ch - 1) / 32 * 32); }

__global__ __launch_bounds__(256) void bss_mean_kernel(const double* x, int ch, int S, double* mean) {
    __shared__ double red[4];
    const int p = blockIdx.x, i = blockIdx.y;
    const double* row = x + ((size_t)p * ch + i) * S;
    double s = 0.0;
    for (int t = threadIdx.x; t < S; t += 256) s += row[t];
    s = block_sum<256>(s, red);
    if (threadIdx.x == 0) mean[(size_t)p * ch + i] = s / S;
}

// Block (p, q) = blockIdx.x (p = / nq, q = % nq), split blockIdx.y of
// `span` samples.  Thread e owns entries e, e + 256, ... of the ch x ch sums.
__global__ __launch_bounds__(256) void bss_cross_kernel(const double* x, const double* mean, int ch, int S,
                                                        const int* lags, int nq, int span, int chunk,
                                                        double* part) {
    const int L = chunk + 1;
    __shared__ double sa[kStage], sb[kStage];
    const int p = blockIdx.x / nq, q = blockIdx.x % nq, split = blockIdx.y;
    const int lag = lags[q];
    const int tn = S - lag;
    const int t_begin = split * span, t_end = min(tn, t_begin + span);
    const double* xp = x + (size_t)p * ch * S;
    const int pairs = ch * ch;
    double acc[kAccPerThread];
#pragma unroll
    for (int r = 0; r < kAccPerThread; ++r) acc[r] = 0.0;
    for (int t0 = t_begin; t0 < t_end; t0 += chunk) {
        const int cw = min(chunk, t_end - t0);
        for (int e = threadIdx.x; e < ch * chunk; e += 256) {
            const int r = e / chunk, c = e % chunk;
            double va = 0.0, vb = 0.0;
            if (c < cw) {
                const double m = mean ? mean[(size_t)p * ch + r] : 0.0;
                va = xp[(size_t)r * S + t0 + c + lag] - m;
                vb = xp[(size_t)r * S + t0 + c] - m;
            }
            sa[r * L + c] = va;
            sb[r * L + c] = vb;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kAccPerThread; ++r) {
            const int e = threadIdx.x + 256 * r;
            if (e >= pairs) break;
            const int i = e / ch, j = e % ch;
            double s = acc[r];
            for (int c = 0; c < cw; ++c) s += sa[i * L + c] * sb[j * L + c];
            acc[r] = s;
        }
        __syncthreads();
    }
    double* out = part + ((size_t)blockIdx.x * gridDim.y + split) * pairs;
#pragma unroll
    for (int r = 0; r < kAccPerThread; ++r) {
        const int e = threadIdx.x + 256 * r;
        if (e >= pairs) break;
        out[e] = acc[r];
    }
}

__global__ void bss_reduce_kernel(const double* part, int ch, int nsplit, const double* denom, int nq, int sym,
                                  double* out, size_t total) {
    const size_t pairs = (size_t)ch * ch;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t blk = e / pairs;  // p nq + q
        const int w = (int)(e % pairs), i = w / ch, j = w % ch;
        const double* pb = part + blk * nsplit * pairs;
        double a = 0.0, b = 0.0;
        for (int s = 0; s < nsplit; ++s) {
            a += pb[(size_t)s * pairs + w];
            if (sym) b += pb[(size_t)s * pairs + (size_t)j * ch + i];
        }
        const double d = denom[blk % nq];
        out[e] = sym ? 0.5 * (a / d + b / d) : a / d;
    }
}

__global__ void bss_whitening_kernel(const double* u, const double* sigma, int ch, double eps, double* W,
                                     double* floor_out) {
    const int p = blockIdx.x;
    const double* up = u + (size_t)p * ch * ch;
    const double* sp = sigma + (size_t)p * ch;
    double smax = 0.0;
    for (int i = 0; i < ch; ++i) smax = fmax(smax, sp[i]);
    const double fl = eps * fmax(smax, 1.0);
    for (int e = threadIdx.x; e < ch * ch; e += blockDim.x) {
        const int i = e / ch, j = e % ch;
        W[(size_t)p * ch * ch + e] = (1.0 / sqrt(fmax(sp[i], fl))) * up[(size_t)j * ch + i];
    }
    if (threadIdx.x == 0 && floor_out) floor_out[p] = fl;
}

// TPT samples per thread (4 at up to 8 channels, 2 at up to 16, else 1),
// their ch inputs each in registers, all loaded before the products;
// op(M) in LDS.  MAXC >= ch bounds the unrolled loops and the LDS tile.
template <int MAXC>
constexpr int bss_apply_tpt() { return MAXC <= 8 ? 4 : (MAXC <= 16 ? 2 : 1); }

template <int MAXC>
__global__ __launch_bounds__(256) void bss_apply_kernel(const double* M, int trans, const double* in,
                                                        const double* mean, int ch, int S, double* out) {
    constexpr int TPT = bss_apply_tpt<MAXC>();
    __shared__ double sm[MAXC * MAXC];
    const int p = blockIdx.x;
    const double* mp = M + (size_t)p * ch * ch;
    for (int e = threadIdx.x; e < ch * ch; e += 256) {
        const int i = e / ch, k = e % ch;
        sm[e] = trans ? mp[(size_t)k * ch + i] : mp[e];
    }
    __syncthreads();
    const double* ip = in + (size_t)p * ch * S;
    double* op = out + (size_t)p * ch * S;
    const int tb = blockIdx.y * 256 * TPT + threadIdx.x;
    double v[TPT][MAXC];
#pragma unroll
    for (int u = 0; u < TPT; ++u) {
        const int t = tb + 256 * u;
#pragma unroll
        for (int k = 0; k < MAXC; ++k)
            v[u][k] = (k < ch && t < S) ? ip[(size_t)k * S + t] - (mean ? mean[(size_t)p * ch + k] : 0.0) : 0.0;
    }
    for (int i = 0; i < ch; ++i) {
#pragma unroll
        for (int u = 0; u < TPT; ++u) {
            const int t = tb + 256 * u;
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < MAXC; ++k)
                if (k < ch) s += sm[i * ch + k] * v[u][k];
            if (t < S) op[(size_t)i * S + t] = s;
        }
    }
}

__global__ void bss_unmixing_kernel(const double* V, const double* W, int ch, double* out) {
    const int p = blockIdx.x;
    const double* vp = V + (size_t)p * ch * ch;
    const double* wp = W + (size_t)p * ch * ch;
    for (int e = threadIdx.x; e < ch * ch; e += blockDim.x) {
        const int i = e / ch, j = e % ch;
        double s = 0.0;
        for (int k = 0; k < ch; ++k) s += vp[(size_t)k * ch + i] * wp[(size_t)k * ch + j];
        out[(size_t)p * ch * ch + e] = s;
    }
}

// channel_phase_coherence (:104-126): 0.5 |<x, x(t+1)>| / |x|^2 +
// 0.5 / (1 + |x(t+1) - x| / |x|), cyclic in t, clamped to [0, 1].
__global__ __launch_bounds__(256) void bss_coherence_kernel(const double* y, int ch, int S, double* coh) {
    __shared__ double red[4];
    const int p = blockIdx.x, i = blockIdx.y;
    const double* row = y + ((size_t)p * ch + i) * S;
    double nrm = 0.0, dd = 0.0, gr = 0.0;
    for (int t = threadIdx.x; t < S; t += 256) {
        const double a = row[t], b = row[t + 1 < S ? t + 1 : 0];
        nrm += a * a;
        dd += a * b;
        gr += (b - a) * (b - a);
    }
    nrm = block_sum<256>(nrm, red);
    dd = block_sum<256>(dd, red);
    gr = block_sum<256>(gr, red);
    if (threadIdx.x == 0) {
        double v = 0.0;
        if (S > 1) {
            const double lock = fabs(dd) / fmax(nrm, 1e-300);
            const double smooth = 1.0 / (1.0 + sqrt(gr / fmax(nrm, 1e-300)));
            v = fmin(fmax(0.5 * lock + 0.5 * smooth, 0.0), 1.0);
        }
        coh[(size_t)p * ch + i] = v;
    }
}

}  // namespace

hipError_t bss_means(const double* x, int count, int ch, int S, double* mean, hipStream_t st) {
    hipLaunchKernelGGL(bss_mean_kernel, dim3(count, ch), dim3(256), 0, st, x, ch, S, mean);
    return hipGetLastError();
}

int bss_splits(int count, int ch, int S, int nq) {
    const int chunks = (S + bss_chunk(ch) - 1) / bss_chunk(ch);
    const int want = (kBssFillBlocks + count * nq - 1) / (count * nq);
    return std::max(1, std::min(chunks, want));
}

hipError_t bss_cross(const double* x, const double* mean, int count, int ch, int S, const int* lags, int nq,
                     int nsplit, double* part, hipStream_t st) {
    const int chunk = bss_chunk(ch);
    const int per = (S + nsplit - 1) / nsplit;
    const int span = (per + chunk - 1) / chunk * chunk;
    hipLaunchKernelGGL(bss_cross_kernel, dim3(count * nq, nsplit), dim3(256), 0, st, x, mean, ch, S, lags, nq, span,
                       chunk, part);
    return hipGetLastError();
}

hipError_t bss_reduce(const double* part, int count, int ch, int nq, int nsplit, const double* denom, bool sym,
                      double* out, hipStream_t st) {
    const size_t total = (size_t)count * nq * ch * ch;
    const int blocks = (int)std::min<size_t>((total + 255) / 256, 4096);
    hipLaunchKernelGGL(bss_reduce_kernel, dim3(blocks > 0 ? blocks : 1), dim3(256), 0, st, part, ch, nsplit, denom,
                       nq, sym ? 1 : 0, out, total);
    return hipGetLastError();
}

hipError_t bss_whitening(const double* u, const double* sigma, int count, int ch, double eps, double* W,
                         double* floor, hipStream_t st) {
    hipLaunchKernelGGL(bss_whitening_kernel, dim3(count), dim3(256), 0, st, u, sigma, ch, eps, W, floor);
    return hipGetLastError();
}

hipError_t bss_apply(const double* M, bool trans, const double* in, const double* mean, int count, int ch, int S,
                     double* out, hipStream_t st) {
    const int tr = trans ? 1 : 0;
    auto grid = [&](int tpt) { return dim3(count, (S + 256 * tpt - 1) / (256 * tpt)); };
    if (ch <= 8)
        hipLaunchKernelGGL(bss_apply_kernel<8>, grid(bss_apply_tpt<8>()), dim3(256), 0, st, M, tr, in, mean, ch, S,
                           out);
    else if (ch <= 16)
        hipLaunchKernelGGL(bss_apply_kernel<16>, grid(bss_apply_tpt<16>()), dim3(256), 0, st, M, tr, in, mean, ch,
                           S, out);
    else if (ch <= 32)
        hipLaunchKernelGGL(bss_apply_kernel<32>, grid(bss_apply_tpt<32>()), dim3(256), 0, st, M, tr, in, mean, ch,
                           S, out);
    else
        hipLaunchKernelGGL(bss_apply_kernel<64>, grid(bss_apply_tpt<64>()), dim3(256), 0, st, M, tr, in, mean, ch,
                           S, out);
    return hipGetLastError();
}

hipError_t bss_unmixing(const double* V, const double* W, int count, int ch, double* out, hipStream_t st) {
    hipLaunchKernelGGL(bss_unmixing_kernel, dim3(count), dim3(256), 0, st, V, W, ch, out);
    return hipGetLastError();
}

hipError_t bss_coherence(const double* y, int count, int ch, int S, double* coh, hipStream_t st) {
    hipLaunchKernelGGL(bss_coherence_kernel, dim3(count, ch), dim3(256), 0, st, y, ch, S, coh);
    return hipGetLastError();
}

}  // namespace lsvd
