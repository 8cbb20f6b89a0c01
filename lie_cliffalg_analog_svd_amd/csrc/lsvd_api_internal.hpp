// lsvd_api_internal.hpp -- handle and plan structures of the host library.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "lsvd_kernels.hpp"
#include "lsvd_workspace.hpp"

namespace lsvd {

struct Bucket {
    int first;           // index of the first item in the device item array
    int count;
    int maxN;
    int maxM;
    size_t arena_bytes;  // largest per-matrix arena in the bucket
    int variant;         // small_svd variant id
    int grid;
    size_t lds_bytes;
    int64_t gws_offset;  // doubles into the plan's global arena
    int64_t gws_stride;
    double cost_sum;     // item_cost summed over the bucket
    double cost_max;
    int consts_off = 0;  // its entries in the plan's constants: the count, and
                         // for the multi-size launch the end of each size class
};

inline uint64_t tune_key(int maxN, size_t arena_bytes, int precision) {
    return ((uint64_t)precision << 56) | ((uint64_t)maxN << 32) | (uint64_t)(arena_bytes & 0xffffffffu);
}

struct Extras;  // lazily created state of the polar / joint / vjp paths

int set_error(struct lsvd_handle_s* h, int code, const std::string& msg);
// LieSvdSmall::solve of `count` square n[b] x n[b] matrices on device pools
// (every n through the small kernels, n <= 4 included); synchronous.
int svd_device_small(struct lsvd_handle_s* h, int count, const int32_t* n, const int64_t* mat_off,
                     const int64_t* s_off, int precision, const double* a, double* u, double* s, double* vt,
                     lsvd_info_t* info);
int hip_fail(struct lsvd_handle_s* h, hipError_t e, const char* what);
// Looks up this device's address of the no-alloc entry point's scratch
// array (once per device; `device` must be current).
hipError_t noalloc_init(int device);

}  // namespace lsvd

struct lsvd_handle_s {
    int device = 0;
    int num_cus = 1;
    std::string arch;
    hipStream_t own_stream = nullptr;
    std::vector<hipStream_t> lanes;              // concurrent bucket lanes, one hardware queue each
    hipStream_t stream = nullptr;
    int ptr_mode = LSVD_PTR_HOST;
    std::string last_error;
    std::map<uint64_t, int> tune;                 // autotune winners: bucket key -> variant
    std::map<uint64_t, lsvd_plan_s*> plan_cache;  // one-item plans of the single-matrix API
    std::vector<std::pair<std::vector<int64_t>, lsvd_plan_s*>> device_plans;  // the library's own batched calls
    int large_min_n = 129;                        // square single solves from this n run chip-wide
    int micro_max_sweeps = 24;                    // LieSvdMicro's sweep limit for n = 3, 4 (test hook)
    bool multi_launch = true;                     // sizes 64..128 of a batch in one launch
    // autotuned plan knobs (lsvd_autotune_batch): padded sizes from
    // multi_min_np up share the multi-size launch (129: none), at most
    // max_lanes concurrent lanes, and bit (np / 16 - 1) of wide_mask runs the
    // size-np bucket with twice the default threads per workgroup
    int multi_min_np = 64;
    // LSVD_PREC_F64_ONESIDED plans: SVD items of padded size <= os_max_np
    // take QR + one-sided Jacobi, the larger ones the scaled Newton-Schulz
    // kernels (autotuned; on the ragged batch every size is faster one-sided:
    // whole plan 5.43 ms at 128 against 7.2-9.1 ms for the smaller splits)
    int os_max_np = 128;
    // precision of the single-matrix entry points' one-item plans
    // (lsvd_set_solve_precision): the reference iteration by default
    int solve_precision = LSVD_PREC_F64;
    int max_lanes = 4;
    int lead_np = 0;  // padded size of the per-size bucket launched first (0: longest estimate first)
    uint32_t wide_mask = 0;
    lsvd_tune_report_t tune_report{};
    // Phase-JADE sweeps run this many bytes of families (working copies and
    // bases) at a time (0: all families together).  About 3/4 of the 256 MB
    // Infinity Cache: 32 families of 64 x 256^2 took 1.25 s unchunked, 1.30 /
    // 1.15 / 1.04 / 1.23 s in chunks of 96 / 128 / 192 / 320 MB.
    size_t jade_chunk_bytes = size_t(192) << 20;
    int lane0_cus = 0;                            // CUs of lane 0 (the longest launch); the other
                                                  // lanes take the rest; 0: every lane on every CU
    lsvd::Extras* extras = nullptr;
    lsvd::Workspace* ws = nullptr;                // temporaries of the entry points (grown, then reused)
};

namespace lsvd {
// Whether a padded size np <= 128 goes to the multi-size launch under the
// handle's knobs, and the per-size variant id otherwise.
inline bool onchip_multi(const lsvd_handle_s* h, int np) {
    return np >= (h->multi_min_np > kV2MultiMinNP ? h->multi_min_np : kV2MultiMinNP) && h->multi_launch;
}
inline int onchip_variant(const lsvd_handle_s* h, int np) {
    return ((h->wide_mask >> (np / 16 - 1)) & 1) && v2_has_wide(np) ? kV2Wide + np : kV2Base + np;
}
}  // namespace lsvd

struct lsvd_plan_s {
    lsvd_handle_t h = nullptr;
    int count = 0;
    int precision = 0;
    int ptr_mode = 0;
    int n_micro = 0;
    int n_items = 0;
    int fallback_first = 0;
    int fallback_grid = 1;
    size_t fallback_lds = 0;
    int n_counters = 0;
    std::vector<lsvd::Bucket> buckets;
    int64_t a_len = 0, u_len = 0, s_len = 0, vt_len = 0;
    lsvd::Item* d_items = nullptr;
    int* d_counters = nullptr;
    int* d_consts = nullptr;
    double* d_gws = nullptr;
    double *d_a = nullptr, *d_u = nullptr, *d_s = nullptr, *d_vt = nullptr;
    lsvd_info_t* d_info = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    lsvd::Pools graph_pools = {};
    std::vector<hipStream_t> side;     // the handle's lanes this plan uses (not owned)
    std::vector<hipEvent_t> ev;        // fork/join events
};

namespace lsvd {
// `tev`, when given, holds 2 * (buckets + 1) timing events: a start/stop
// pair recorded on the lane stream around the micro launch (pair 0) and
// around each bucket's launch (pair 1 + i).
hipError_t plan_enqueue(lsvd_plan_t p, const Pools& pools, hipStream_t st, const hipEvent_t* tev = nullptr);
}

// Frees the handle's polar / joint / vjp state (defined with those paths).
void lsvd_release_extras(lsvd_handle_t h);
