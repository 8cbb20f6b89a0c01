// metrics.hip -- metrics::compute (src/lib.rs:149-194) over a whole batch
// of factorizations on the device, so that the residual checks of a
// 12288-matrix batch cover every matrix instead of a host-side sample.
//
// Per matrix (m x n, k = min(m, n); U m x k, sigma k, Vt k x n, row-major in
// packed pools):
//   rel_recon = ||U diag(sigma) Vt - A||_F / max(||A||_F, 1e-300)
//   orth_u    = ||U^T U - I_k||_F,   orth_v = ||Vt Vt^T - I_k||_F
//   sigma_max_rel / sigma_tail_rel against a reference spectrum, both sorted
//   descending: max |got - want| / max(|want|, 1e-300) over all i, and over
//   i >= floor(3 m / 4) (the reference's tail start, n = a.nrows()).
// The reference's identities are n x n for its square factors; thin factors
// take the k x k identity.  One workgroup per matrix (grid-stride over the
// batch); entries are accumulated in the reference's order, (U_il s_l) Vt_lj
// summed over l.  HBM/L2-bound: each matrix's factors are read from L2
// k times over (a few MB for the largest), no tiles.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "lsvd_api_internal.hpp"
#include "lsvd_device.hpp"

namespace lsvd {

constexpr int kMetricsNT = 256;
constexpr int kMetricsMaxK = 1024;

__global__ __launch_bounds__(kMetricsNT) void metrics_kernel(const Item* items, int count, const double* a,
                                                               const double* u, const double* s, const double* vt,
                                                               const double* sref, lsvd_metrics_t* out) {
    __shared__ double red[kMetricsNT / kWave];
    __shared__ double got[kMetricsMaxK], want[kMetricsMaxK];
    const int tid = tix();
    for (int b = blockIdx.x; b < count; b += gridDim.x) {
        const Item it = items[b];
        const int m = it.m, n = it.n, k = m < n ? m : n;
        const double* A = a + it.a_off;
        const double* U = u + it.u_off;
        const double* S = s + it.s_off;
        const double* V = vt + it.vt_off;
        double ra = 0.0, rd = 0.0;
        for (int e = tid; e < m * n; e += kMetricsNT) {
            const int i = e / n, j = e % n;
            double acc = 0.0;
            for (int l = 0; l < k; ++l) acc += (U[(size_t)i * k + l] * S[l]) * V[(size_t)l * n + j];
            const double x = A[e];
            const double d = acc - x;
            rd += d * d;
            ra += x * x;
        }
        double ou = 0.0, ov = 0.0;
        for (int e = tid; e < k * k; e += kMetricsNT) {
            const int p = e / k, q = e % k;
            double gu = 0.0, gv = 0.0;
            for (int i = 0; i < m; ++i) gu += U[(size_t)i * k + p] * U[(size_t)i * k + q];
            for (int j = 0; j < n; ++j) gv += V[(size_t)p * n + j] * V[(size_t)q * n + j];
            const double id = p == q ? 1.0 : 0.0;
            ou += (gu - id) * (gu - id);
            ov += (gv - id) * (gv - id);
        }
        rd = block_sum<kMetricsNT>(rd, red);
        ra = block_sum<kMetricsNT>(ra, red);
        ou = block_sum<kMetricsNT>(ou, red);
        ov = block_sum<kMetricsNT>(ov, red);
        double smr = NAN, stl = NAN;
        if (sref) {
            const double* W = sref + it.s_off;
            // descending rank sort of both spectra (ties by index)
            for (int i = tid; i < k; i += kMetricsNT) {
                const double x = S[i], y = W[i];
                int rx = 0, ry = 0;
                for (int j = 0; j < k; ++j) {
                    const double xj = S[j], yj = W[j];
                    rx += (xj > x) || (xj == x && j < i);
                    ry += (yj > y) || (yj == y && j < i);
                }
                got[rx] = x;
                want[ry] = y;
            }
            __syncthreads();
            const int tail = (m * 3) / 4;
            double emax = 0.0, etail = 0.0;
            for (int i = tid; i < k; i += kMetricsNT) {
                const double e = fabs(got[i] - want[i]) / fmax(fabs(want[i]), 1e-300);
                emax = fmax(emax, e);
                if (i >= tail) etail = fmax(etail, e);
            }
            // block max through the sum helper's slots (values are >= 0)
            emax = wave_max(emax);
            etail = wave_max(etail);
            __syncthreads();
            if ((tid & (kWave - 1)) == 0) {
                got[tid / kWave] = emax;
                want[tid / kWave] = etail;
            }
            __syncthreads();
            smr = 0.0;
            stl = 0.0;
            for (int w = 0; w < kMetricsNT / kWave; ++w) {
                smr = fmax(smr, got[w]);
                stl = fmax(stl, want[w]);
            }
            __syncthreads();
        }
        if (tid == 0) {
            lsvd_metrics_t r;
            r.rel_recon = sqrt(rd) / fmax(sqrt(ra), 1e-300);
            r.orth_u = sqrt(ou);
            r.orth_v = sqrt(ov);
            r.sigma_max_rel = smr;
            r.sigma_tail_rel = stl;
            out[it.idx] = r;
        }
    }
}

}  // namespace lsvd

using namespace lsvd;

#define MCHK(expr)                                          \
    do {                                                    \
        hipError_t e_ = (expr);                             \
        if (e_ != hipSuccess) return hip_fail(h, e_, #expr); \
    } while (0)

extern "C" int lsvd_metrics_batched_f64(lsvd_handle_t h, int count, const int32_t* m, const int32_t* n,
                                        const int64_t* a_off, const int64_t* u_off, const int64_t* s_off,
                                        const int64_t* vt_off, const double* a, const double* u, const double* s,
                                        const double* vt, const double* sigma_ref, lsvd_metrics_t* out) {
    if (!h) return LSVD_E_ARG;
    if (count < 0) return set_error(h, LSVD_E_SHAPE, "negative count");
    if (count == 0) return LSVD_OK;
    if (!m || !n || !a_off || !u_off || !s_off || !vt_off || !a || !u || !s || !vt || !out)
        return set_error(h, LSVD_E_ARG, "null pointer");
    std::vector<Item> items(count);
    int64_t a_len = 1, u_len = 1, s_len = 1, vt_len = 1;
    for (int b = 0; b < count; ++b) {
        const int mb = m[b], nb = n[b], k = mb < nb ? mb : nb;
        if (mb <= 0 || nb <= 0) return set_error(h, LSVD_E_SHAPE, "matrix " + std::to_string(b) + ": empty");
        if (k > kMetricsMaxK)
            return set_error(h, LSVD_E_LIMIT, "matrix " + std::to_string(b) + ": min(m, n) > 1024");
        items[b] = Item{mb, nb, b, kModeSvd, a_off[b], u_off[b], s_off[b], vt_off[b]};
        a_len = std::max<int64_t>(a_len, a_off[b] + (int64_t)mb * nb);
        u_len = std::max<int64_t>(u_len, u_off[b] + (int64_t)mb * k);
        s_len = std::max<int64_t>(s_len, s_off[b] + k);
        vt_len = std::max<int64_t>(vt_len, vt_off[b] + (int64_t)k * nb);
    }
    MCHK(hipSetDevice(h->device));
    hipStream_t st = h->stream;
    WsScope ws_scope(h);
    WsMem bitems, ba, bu, bs, bvt, bref, bout;
    MCHK(bitems.alloc(sizeof(Item) * count));
    MCHK(hipMemcpyAsync(bitems.p, items.data(), sizeof(Item) * count, hipMemcpyHostToDevice, st));
    const bool host = h->ptr_mode == LSVD_PTR_HOST;
    const double *da = a, *du = u, *ds = s, *dvt = vt, *dref = sigma_ref;
    lsvd_metrics_t* dout = out;
    if (host) {
        auto put = [&](WsMem& buf, const double* src, int64_t len, const double** dst) -> hipError_t {
            hipError_t e = buf.alloc(sizeof(double) * len);
            if (e == hipSuccess) e = hipMemcpyAsync(buf.p, src, sizeof(double) * len, hipMemcpyHostToDevice, st);
            *dst = static_cast<const double*>(buf.p);
            return e;
        };
        MCHK(put(ba, a, a_len, &da));
        MCHK(put(bu, u, u_len, &du));
        MCHK(put(bs, s, s_len, &ds));
        MCHK(put(bvt, vt, vt_len, &dvt));
        if (sigma_ref) MCHK(put(bref, sigma_ref, s_len, &dref));
        MCHK(bout.alloc(sizeof(lsvd_metrics_t) * count));
        dout = static_cast<lsvd_metrics_t*>(bout.p);
    }
    const int grid = std::min(count, 16 * h->num_cus);
    hipLaunchKernelGGL(metrics_kernel, dim3(grid), dim3(kMetricsNT), 0, st, static_cast<const Item*>(bitems.p),
                       count, da, du, ds, dvt, dref, dout);
    MCHK(hipGetLastError());
    if (host) MCHK(hipMemcpyAsync(out, dout, sizeof(lsvd_metrics_t) * count, hipMemcpyDeviceToHost, st));
    MCHK(hipStreamSynchronize(st));  // the item list and host copies live in the workspace
    return LSVD_OK;
}
