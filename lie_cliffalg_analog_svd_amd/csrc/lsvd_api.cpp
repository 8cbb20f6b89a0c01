// lsvd_api.cpp -- host side of the C ABI declared in include/lie_svd_amd.h.
//
// A handle owns a HIP stream, its lanes and the autotuner's choices.  Every
// solve is expressed as a batch plan: a list of Items sorted longest-first
// and split into buckets by the tall-ified shape (M x N, M >= N).  Each
// bucket is one persistent launch sized for that bucket: the reference
// iteration's kernels (padded sizes 64..128 share one launch) or, for
// LSVD_PREC_F64_ONESIDED, the QR + one-sided Jacobi kernels (padded sizes
// 80..128 share one launch); square n <= 4 items go to the micro kernel,
// whose escalations are drained by a last small_svd launch.  A plan whose buckets fit one lane replays a captured hipGraph; a
// plan spread over several lanes is enqueued on them directly (see
// lsvd_batch_plan_execute for the measured reason).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "lsvd_api_internal.hpp"
#include "lsvd_large.hpp"

using namespace lsvd;

#define LSVD_VERSION "0.1.0"

namespace lsvd {

int set_error(lsvd_handle_t h, int code, const std::string& msg) {
    if (h) h->last_error = msg;
    return code;
}

int hip_fail(lsvd_handle_t h, hipError_t e, const char* what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return set_error(h, LSVD_E_HIP, m);
}

// ---- bucket selection --------------------------------------------------

static size_t lds_budget_bytes() { return 160 * 1024 - 8 * 1024; }  // minus static LDS headroom

// N class boundaries; each class gets its own persistent launch.
static int n_class(int N) {
    static const int bounds[] = {8, 16, 24, 32, 48, 64, 96, 128, 256, 512, 1024};
    for (int i = 0; i < (int)(sizeof(bounds) / sizeof(bounds[0])); ++i)
        if (N <= bounds[i]) return i;
    return -1;
}

// Default variant for a bucket: the smallest MAXN that holds N, with the
// arena in LDS when the bucket's largest arena fits there.
static int default_variant(int maxN, size_t arena_bytes) {
    const bool fits = arena_bytes <= lds_budget_bytes();
    if (maxN <= 16 && fits) return 0;
    if (maxN <= 32 && fits) return 1;
    if (maxN <= 64 && fits) return 2;
    if (maxN <= 128) return 3;
    if (maxN <= 1024) return 4;
    return -1;
}

static bool variant_accepts(int vid, int maxN, size_t arena_bytes) {
    int nv = 0;
    const SmallVariant* vs = small_variants(&nv);
    for (int i = 0; i < nv; ++i)
        if (vs[i].id == vid) return maxN <= vs[i].maxn && (!vs[i].lds || arena_bytes <= lds_budget_bytes());
    return false;
}

static const SmallVariant* find_variant(int vid) {
    int nv = 0;
    const SmallVariant* vs = small_variants(&nv);
    for (int i = 0; i < nv; ++i)
        if (vs[i].id == vid) return &vs[i];
    return nullptr;
}

// Work estimate used for longest-first ordering (flops up to a constant):
// QR + compose ~ 4 M N^2, NS ~ 4 k N^3 with k ~ 30, Jacobi ~ 6 S N^3, S ~ 8.
// Relative cost used to order the queues (largest first): about 42
// Newton-Schulz steps of two M x N x N products, Jacobi sweeps of ~12 N^3.
static double item_cost(int M, int N) {
    const double n3 = (double)N * N * N;
    return 86.0 * M * (double)N * N + 48.0 * n3;
}

// Sets the launch shape of a bucket for `variant`; returns the global arena
// doubles it needs (0 for LDS variants).
static size_t configure_bucket(lsvd_handle_t h, Bucket& bk, int variant) {
    if (variant >= kOsMulti) {  // one-sided sizes 80..128 in one launch: LDS and slot for the largest
        const int np = variant - kOsMulti;
        bk.variant = variant;
        const size_t lds = os_lds_bytes(np, bk.maxM);
        const int per_cu = onesided_multi_blocks_per_cu(lds);
        bk.grid = std::max(1, std::min(bk.count, per_cu * h->num_cus));
        bk.lds_bytes = lds;
        bk.gws_stride = os_slot_doubles(np, bk.maxM);
        return (size_t)bk.gws_stride * bk.grid;
    }
    if (variant >= kOsBase) {  // QR + one-sided Jacobi: LDS and reflector slot sized by the largest M
        const int np = variant - kOsBase;
        bk.variant = variant;
        const size_t lds = os_lds_bytes(np, bk.maxM);
        const int per_cu = onesided_blocks_per_cu(np, lds);
        bk.grid = std::max(1, std::min(bk.count, per_cu * h->num_cus));
        bk.lds_bytes = lds;
        bk.gws_stride = os_slot_doubles(np, bk.maxM);
        return (size_t)bk.gws_stride * bk.grid;
    }
    if (variant >= kV2Multi && variant < kV2Wide) {  // one launch over sizes 64..128: LDS and slots for the largest item
        const int np = variant - kV2Multi;
        bk.variant = variant;
        const size_t lds = v2_lds_bytes(np, bk.maxM);
        const int per_cu = small2_multi_blocks_per_cu(lds);
        const int cus = h->lane0_cus > 0 && h->lane0_cus < h->num_cus ? h->lane0_cus : h->num_cus;
        bk.grid = std::max(1, std::min(bk.count, per_cu * cus));
        bk.lds_bytes = lds;
        bk.gws_stride = v2_slot_doubles(np, bk.maxM);
        return (size_t)bk.gws_stride * bk.grid;
    }
    if (variant >= kV2Base) {  // on-chip kernel: LDS sized by the bucket's largest M
        const bool wide = variant >= kV2Wide;
        const int np = variant - (wide ? kV2Wide : kV2Base);
        bk.variant = variant;
        const size_t lds = v2_lds_bytes(np, bk.maxM);
        const int per_cu = small2_blocks_per_cu(np, wide, lds);
        bk.grid = std::max(1, std::min(bk.count, per_cu * h->num_cus));
        bk.lds_bytes = lds;
        bk.gws_stride = v2_slot_doubles(np, bk.maxM);
        return (size_t)bk.gws_stride * bk.grid;
    }
    const SmallVariant* v = find_variant(variant);
    bk.variant = variant;
    const size_t lds = v->lds ? bk.arena_bytes : 0;
    const int per_cu = small_variant_blocks_per_cu(variant, lds);
    bk.grid = std::max(1, std::min(bk.count, per_cu * h->num_cus));
    bk.lds_bytes = lds;
    bk.gws_stride = v->lds ? 0 : (int64_t)((bk.arena_bytes / 8 + 31) / 32 * 32);
    return v->lds ? 0 : (size_t)bk.gws_stride * bk.grid;
}

}  // namespace lsvd

// ---- plan -----------------------------------------------------------------

static int plan_build(lsvd_handle_t h, int count, const int32_t* m, const int32_t* n, const int64_t* a_off,
                      const int64_t* u_off, const int64_t* s_off, const int64_t* vt_off, const int32_t* modes,
                      int precision, lsvd_plan_t* out) {
    if (!h || !out) return LSVD_E_ARG;
    if (count < 0) return set_error(h, LSVD_E_ARG, "negative count");
    if (precision != LSVD_PREC_F64 && precision != LSVD_PREC_F64_SCALED && precision != LSVD_PREC_F64_ONESIDED)
        return set_error(h, LSVD_E_ARG,
                         "batch plans compute in FP64: precision must be LSVD_PREC_F64, LSVD_PREC_F64_SCALED or "
                         "LSVD_PREC_F64_ONESIDED");
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(h, e, "hipSetDevice");

    auto* p = new lsvd_plan_s();
    p->h = h;
    p->count = count;
    p->precision = precision;

    std::vector<Item> micro;
    struct Tmp {
        Item it;
        int M, N;
        double cost;
    };
    std::map<int, std::vector<Tmp>> by_class;
    int64_t a_end = 0, u_end = 0, s_end = 0, vt_end = 0;
    for (int b = 0; b < count; ++b) {
        const int mb = m[b], nb = n[b];
        const int mode = modes ? modes[b] : kModeSvd;
        if (mb < 0 || nb < 0 || (mode != kModeSvd && mb != nb)) {
            delete p;
            return set_error(h, LSVD_E_SHAPE, "matrix " + std::to_string(b) + ": bad shape");
        }
        const int k = std::min(mb, nb);
        Item it{};
        it.m = mb;
        it.n = nb;
        it.idx = b;
        it.mode = mode;
        it.a_off = a_off[b];
        it.u_off = u_off[b];
        it.s_off = s_off[b];
        it.vt_off = vt_off ? vt_off[b] : 0;
        a_end = std::max<int64_t>(a_end, it.a_off + (int64_t)mb * nb);
        u_end = std::max<int64_t>(u_end, it.u_off + (int64_t)(mode == kModeEigh ? nb * nb : mb * k));
        s_end = std::max<int64_t>(s_end, it.s_off + k);
        if (mode != kModeEigh) vt_end = std::max<int64_t>(vt_end, it.vt_off + (int64_t)k * nb);
        if (mb == 0 || nb == 0) {
            if (mode == kModeSvd) {  // LieSvdSmall::solve_rectangular asserts non-empty (:343-346)
                delete p;
                return set_error(h, LSVD_E_SHAPE, "LieSvdSmall::solve_rectangular: empty input");
            }
            continue;  // LieSvdMicro / eigh on 0x0: nothing to do
        }
        if (mode == kModeMicro) {
            if (nb > 4) {
                delete p;
                return set_error(h, LSVD_E_SHAPE, "LieSvdMicro only supports N <= 4");
            }
            micro.push_back(it);
            continue;
        }
        const int M = std::max(mb, nb), N = std::min(mb, nb);
        int key;
        if (N <= 128 && M <= 128 && precision == LSVD_PREC_F64_ONESIDED && mode != kModeEigh &&
            (N + 15) / 16 * 16 <= h->os_max_np) {
            const int np = (N + 15) / 16 * 16;
            // QR + one-sided Jacobi: one launch per padded size below 80, one
            // launch over the larger sizes (they share the launch shape)
            key = np >= kOsMultiMinNP ? 3500 : 3000 + np;
        } else if (N <= 128 && M <= 128) {
            const int np = (N + 15) / 16 * 16;
            // on-chip kernels: one launch per padded size below 64, one
            // launch over all the larger sizes (a queue sorted longest first)
            key = onchip_multi(h, np) ? 2000 : 1000 + np;
        } else {
            const int cls = n_class(N);
            if (cls < 0) {
                delete p;
                return set_error(h, LSVD_E_LIMIT, "matrix " + std::to_string(b) + ": min(m, n) > 1024");
            }
            // split classes further by whether the arena is large (tall M)
            const size_t arena = small_arena_doubles(M, N) * 8;
            key = cls * 2 + (arena > lds_budget_bytes() ? 1 : 0);
        }
        by_class[key].push_back({it, M, N, item_cost(M, N)});
    }
    p->a_len = a_end;
    p->u_len = u_end;
    p->s_len = s_end;
    p->vt_len = vt_end;

    // Device layout: [micro items][bucket items...][fallback items]
    std::vector<Item> all;
    all.insert(all.end(), micro.begin(), micro.end());
    p->n_micro = (int)micro.size();
    std::vector<int> consts;  // per bucket: its count (and the multi launch's class ends)
    for (auto& kv : by_class) {
        auto& v = kv.second;
        std::stable_sort(v.begin(), v.end(), [](const Tmp& x, const Tmp& y) { return x.cost > y.cost; });
        if (kv.first == 2000 || kv.first == 3500)  // the multi-size launches drain their size classes largest first
            std::stable_sort(v.begin(), v.end(),
                             [](const Tmp& x, const Tmp& y) { return (x.N + 15) / 16 > (y.N + 15) / 16; });
        Bucket bk{};
        bk.first = (int)all.size();
        bk.count = (int)v.size();
        int maxN = 0, maxM = 0;
        size_t arena = 0;
        for (auto& t : v) {
            all.push_back(t.it);
            bk.cost_sum += t.cost;
            bk.cost_max = std::max(bk.cost_max, t.cost);
            maxN = std::max(maxN, t.N);
            maxM = std::max(maxM, t.M);
            arena = std::max(arena, small_arena_doubles(t.M, t.N) * 8);
        }
        bk.maxN = maxN;
        bk.maxM = maxM;
        bk.arena_bytes = arena;
        bk.consts_off = (int)consts.size();
        consts.push_back(bk.count);
        if (kv.first == 2000 || kv.first == 3500) {  // class ends 128, 112, 96, 80[, 64] (relative to bk.first)
            for (int np = 128; np >= (kv.first == 2000 ? kV2MultiMinNP : kOsMultiMinNP); np -= 16) {
                int end = 0;
                while (end < (int)v.size() && (v[end].N + 15) / 16 * 16 >= np) ++end;
                consts.push_back(end);
            }
        }
        int vid;
        if (kv.first == 3500) {
            vid = kOsMulti + (maxN + 15) / 16 * 16;
        } else if (kv.first >= 3000) {
            vid = kOsBase + (kv.first - 3000);
        } else if (kv.first == 2000) {
            vid = kV2Multi + (maxN + 15) / 16 * 16;
        } else if (kv.first >= 1000) {
            const int np = kv.first - 1000;
            vid = onchip_variant(h, np);
        } else {
            vid = default_variant(maxN, arena);
            auto tk = h->tune.find(tune_key(maxN, arena, precision));
            if (tk != h->tune.end() && variant_accepts(tk->second, maxN, arena)) vid = tk->second;
        }
        if (vid < 0) {
            delete p;
            return set_error(h, LSVD_E_LIMIT, "no kernel variant for N=" + std::to_string(maxN));
        }
        bk.variant = vid;
        p->buckets.push_back(bk);
    }
    p->fallback_first = (int)all.size();
    p->n_items = (int)all.size() + p->n_micro;  // room for fallbacks
    all.resize(p->n_items);

    // counters: [bucket counters][fallback counter][fallback count]
    const int nb = (int)p->buckets.size();
    p->n_counters = nb + 2;
    consts.push_back(0);

    // global arena slots for non-LDS variants, and launch grids
    size_t gws_doubles = 0;
    for (auto& bk : p->buckets) {
        bk.gws_offset = (int64_t)gws_doubles;
        gws_doubles += configure_bucket(h, bk, bk.variant);
    }
    // fallback launch: micro escalations, n <= 4 square -> variant 0.  It
    // runs after the join and its queue is usually empty, so its grid is
    // at most one workgroup per CU (a full-occupancy grid of workgroups
    // that find nothing to do cost ~45 us per execution)
    {
        const size_t arena = small_arena_doubles(4, 4) * 8;
        const int per_cu = small_variant_blocks_per_cu(0, arena);
        p->fallback_grid = std::max(1, std::min({std::max(p->n_micro, 1), per_cu * h->num_cus, h->num_cus}));
        p->fallback_lds = arena;
    }

#define PCHECK(expr)                                        \
    do {                                                    \
        hipError_t e_ = (expr);                             \
        if (e_ != hipSuccess) {                             \
            int rc_ = hip_fail(h, e_, #expr);               \
            lsvd_batch_plan_destroy(p);                     \
            return rc_;                                     \
        }                                                   \
    } while (0)

    PCHECK(hipMalloc(&p->d_items, sizeof(Item) * std::max(1, p->n_items)));
    if (!all.empty())
        PCHECK(hipMemcpy(p->d_items, all.data(), sizeof(Item) * all.size(), hipMemcpyHostToDevice));
    PCHECK(hipMalloc(&p->d_counters, sizeof(int) * p->n_counters));
    PCHECK(hipMalloc(&p->d_consts, sizeof(int) * consts.size()));
    PCHECK(hipMemcpy(p->d_consts, consts.data(), sizeof(int) * consts.size(), hipMemcpyHostToDevice));
    if (gws_doubles) PCHECK(hipMalloc(&p->d_gws, sizeof(double) * gws_doubles));
    if (h->ptr_mode == LSVD_PTR_HOST) {
        PCHECK(hipMalloc(&p->d_a, sizeof(double) * std::max<int64_t>(1, p->a_len)));
        PCHECK(hipMalloc(&p->d_u, sizeof(double) * std::max<int64_t>(1, p->u_len)));
        PCHECK(hipMalloc(&p->d_s, sizeof(double) * std::max<int64_t>(1, p->s_len)));
        PCHECK(hipMalloc(&p->d_vt, sizeof(double) * std::max<int64_t>(1, p->vt_len)));
        PCHECK(hipMalloc(&p->d_info, sizeof(lsvd_info_t) * std::max(1, count)));
    }
#undef PCHECK
    p->ptr_mode = h->ptr_mode;
    *out = p;
    return LSVD_OK;
}

namespace lsvd {

// Lanes (streams) and fork/join events for the concurrent buckets.
// Created outside any graph capture (plan_run_device calls this first).
// The lanes belong to the handle and are created with a CU mask covering
// every CU: such a stream gets a hardware queue of its own, whereas plain
// streams share the process's few queues in creation order, and two lanes
// on one queue run their buckets one after the other.
constexpr int kPlanLanes = 4;

static int plan_lanes(lsvd_plan_t p) {
    const int cap = std::max(1, std::min(kPlanLanes, p->h->max_lanes));
    return std::min(cap, (int)p->buckets.size() + (p->n_micro > 0 ? 1 : 0));
}

hipError_t plan_ensure_streams(lsvd_plan_t p) {
    const int lanes = plan_lanes(p);
    lsvd_handle_t h = p->h;
    while ((int)h->lanes.size() < lanes) {
        uint32_t mask[32];
        for (uint32_t& w : mask) w = 0xffffffffu;
        const uint32_t words = (uint32_t)std::min((h->num_cus + 31) / 32, 32);
        if (h->lane0_cus > 0 && h->lane0_cus < h->num_cus) {  // lane 0: CUs [0, lane0_cus), the others the rest
            const bool first = h->lanes.empty();
            for (int c = 0; c < 32 * 32; ++c) {
                const bool on = c < h->num_cus && (first ? c < h->lane0_cus : c >= h->lane0_cus);
                if (on) mask[c / 32] |= 1u << (c % 32);
                else mask[c / 32] &= ~(1u << (c % 32));
            }
        }
        hipStream_t s2;
        hipError_t e = hipExtStreamCreateWithCUMask(&s2, words, mask);
        if (e != hipSuccess) e = hipStreamCreateWithFlags(&s2, hipStreamNonBlocking);
        if (e != hipSuccess) return e;
        h->lanes.push_back(s2);
    }
    p->side.assign(h->lanes.begin(), h->lanes.begin() + lanes);
    while ((int)p->ev.size() < lanes + 1) {
        hipEvent_t ev;
        hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e != hipSuccess) return e;
        p->ev.push_back(ev);
    }
    return hipSuccess;
}

// Padded size of a per-size bucket (0 for the multi-size launches and the
// variants above 128).
static int bucket_np(const Bucket& bk) {
    const int v = bk.variant;
    if (v >= kOsMulti) return 0;
    if (v >= kOsBase) return v - kOsBase;
    if (v >= kV2Wide) return v - kV2Wide;
    if (v >= kV2Multi) return 0;
    if (v >= kV2Base) return v - kV2Base;
    return 0;
}

// Enqueues the whole plan on `st` with the given device pools.  The
// buckets (and the micro launch) are spread over up to four concurrent
// lanes (fork/join with events, which a hipGraph capture turns into
// parallel branches): longest estimated bucket first, each onto the lane
// with the least estimated work, so a lane never queues a long bucket
// behind another one while the CUs a large-N bucket leaves idle take the
// small-N buckets.  The handle's lead_np (autotuned) moves one per-size
// bucket to the front: launched first, its workgroups take CUs before the
// persistent multi-size launch fills them.  The micro escalations drain
// after the join.
hipError_t plan_enqueue(lsvd_plan_t p, const Pools& pools, hipStream_t st, const hipEvent_t* tev) {
    hipError_t e = launch_reset_counters(p->d_counters, p->n_counters, st);
    if (e != hipSuccess) return e;
    const int nb = (int)p->buckets.size();
    int* fb_count = p->d_counters + nb + 1;
    e = plan_ensure_streams(p);
    if (e != hipSuccess) return e;
    const int lanes = plan_lanes(p);
    std::vector<int> order(nb);
    for (int i = 0; i < nb; ++i) order[i] = i;
    auto est = [&](int i) {
        const Bucket& bk = p->buckets[i];
        return bk.cost_sum / std::max(bk.grid, 1) + bk.cost_max;
    };
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return est(x) > est(y); });
    if (p->h->lead_np > 0) {
        auto it = std::find_if(order.begin(), order.end(),
                               [&](int i) { return bucket_np(p->buckets[i]) == p->h->lead_np; });
        if (it != order.end()) std::rotate(order.begin(), it, it + 1);
    }
    std::vector<double> load(std::max(lanes, 1), 0.0);
    auto least = [&]() { return (int)(std::min_element(load.begin(), load.end()) - load.begin()); };
    e = hipEventRecord(p->ev[0], st);  // fork point
    if (e != hipSuccess) return e;
    for (int l = 0; l < lanes; ++l) {
        e = hipStreamWaitEvent(p->side[l], p->ev[0], 0);
        if (e != hipSuccess) return e;
    }
    for (int i : order) {
        const Bucket& bk = p->buckets[i];
        const int l = least();
        load[l] += est(i);
        SmallLaunch L{};
        L.grid = bk.grid;
        L.lds_bytes = bk.lds_bytes;
        L.items = p->d_items + bk.first;
        L.count_ptr = p->d_consts + bk.consts_off;
        L.counter = p->d_counters + i;
        L.pools = pools;
        L.gws = p->d_gws ? p->d_gws + bk.gws_offset : nullptr;
        L.gws_stride = bk.gws_stride;
        L.ns_scaled = p->precision != LSVD_PREC_F64;
        if (tev && (e = hipEventRecord(tev[2 * (i + 1)], p->side[l])) != hipSuccess) return e;
        e = launch_small_variant(bk.variant, L, p->side[l]);
        if (e != hipSuccess) return e;
        if (tev && (e = hipEventRecord(tev[2 * (i + 1) + 1], p->side[l])) != hipSuccess) return e;
    }
    if (p->n_micro > 0) {
        const int l = least();
        MicroLaunch ml{p->d_items, p->n_micro, pools, p->d_items + p->fallback_first, fb_count,
                       p->h->micro_max_sweeps};
        if (tev && (e = hipEventRecord(tev[0], p->side[l])) != hipSuccess) return e;
        e = launch_micro(ml, p->side[l]);
        if (e != hipSuccess) return e;
        if (tev && (e = hipEventRecord(tev[1], p->side[l])) != hipSuccess) return e;
    }
    for (int l = 0; l < lanes; ++l) {  // join
        e = hipEventRecord(p->ev[l + 1], p->side[l]);
        if (e != hipSuccess) return e;
        e = hipStreamWaitEvent(st, p->ev[l + 1], 0);
        if (e != hipSuccess) return e;
    }
    if (p->n_micro > 0) {
        SmallLaunch L{};
        L.grid = p->fallback_grid;
        L.lds_bytes = p->fallback_lds;
        L.items = p->d_items + p->fallback_first;
        L.count_ptr = fb_count;
        L.counter = p->d_counters + nb;
        L.pools = pools;
        e = launch_small_variant(0, L, st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace lsvd

static int plan_run_device(lsvd_plan_t p, const Pools& pools, bool use_graph) {
    lsvd_handle_t h = p->h;
    hipStream_t st = h->stream;
    hipError_t e;
    if (!use_graph) {
        e = plan_enqueue(p, pools, st);
        return e == hipSuccess ? LSVD_OK : hip_fail(h, e, "kernel launch");
    }
    const bool same = p->graph_exec && p->graph_pools.a == pools.a && p->graph_pools.u == pools.u &&
                      p->graph_pools.s == pools.s && p->graph_pools.vt == pools.vt &&
                      p->graph_pools.info == pools.info;
    if (!same) {
        if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
        if (p->graph) (void)hipGraphDestroy(p->graph);
        p->graph_exec = nullptr;
        p->graph = nullptr;
        e = plan_ensure_streams(p);
        if (e != hipSuccess) return hip_fail(h, e, "hipStreamCreate");
        e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return hip_fail(h, e, "hipStreamBeginCapture");
        hipError_t le = plan_enqueue(p, pools, st);
        e = hipStreamEndCapture(st, &p->graph);
        if (le != hipSuccess) return hip_fail(h, le, "kernel launch (capture)");
        if (e != hipSuccess) return hip_fail(h, e, "hipStreamEndCapture");
        e = hipGraphInstantiate(&p->graph_exec, p->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) return hip_fail(h, e, "hipGraphInstantiate");
        p->graph_pools = pools;
    }
    e = hipGraphLaunch(p->graph_exec, st);
    return e == hipSuccess ? LSVD_OK : hip_fail(h, e, "hipGraphLaunch");
}

static int plan_execute(lsvd_plan_t p, const double* a_pool, double* u_pool, double* s_pool, double* vt_pool,
                        lsvd_info_t* info, bool use_graph) {
    if (!p) return LSVD_E_ARG;
    lsvd_handle_t h = p->h;
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(h, e, "hipSetDevice");
    if (p->count > 0 && (!a_pool || !u_pool || !s_pool || (!vt_pool && p->vt_len > 0)))
        return set_error(h, LSVD_E_ARG, "null pool pointer");
    if (p->ptr_mode == LSVD_PTR_DEVICE) {
        Pools pools{a_pool, u_pool, s_pool, vt_pool, info};
        return plan_run_device(p, pools, use_graph);
    }
    hipStream_t st = h->stream;
#define ECHECK(expr)                                       \
    do {                                                   \
        hipError_t e_ = (expr);                            \
        if (e_ != hipSuccess) return hip_fail(h, e_, #expr); \
    } while (0)
    if (p->a_len) ECHECK(hipMemcpyAsync(p->d_a, a_pool, sizeof(double) * p->a_len, hipMemcpyHostToDevice, st));
    Pools pools{p->d_a, p->d_u, p->d_s, p->d_vt, info ? p->d_info : nullptr};
    int rc = plan_run_device(p, pools, use_graph);
    if (rc != LSVD_OK) return rc;
    if (p->u_len) ECHECK(hipMemcpyAsync(u_pool, p->d_u, sizeof(double) * p->u_len, hipMemcpyDeviceToHost, st));
    if (p->s_len) ECHECK(hipMemcpyAsync(s_pool, p->d_s, sizeof(double) * p->s_len, hipMemcpyDeviceToHost, st));
    if (p->vt_len) ECHECK(hipMemcpyAsync(vt_pool, p->d_vt, sizeof(double) * p->vt_len, hipMemcpyDeviceToHost, st));
    if (info && p->count)
        ECHECK(hipMemcpyAsync(info, p->d_info, sizeof(lsvd_info_t) * p->count, hipMemcpyDeviceToHost, st));
    ECHECK(hipStreamSynchronize(st));
#undef ECHECK
    return LSVD_OK;
}

namespace lsvd {
// Device-pointer plans of the library's own batched calls (the BSS
// covariances, HOSVD mode Grams, polish cores, vjp forwards), cached per
// handle by their exact item list so that a repeated call builds and
// allocates nothing.
static lsvd_plan_t cached_device_plan(lsvd_handle_t h, int count, const int32_t* m, const int32_t* n,
                                      const int64_t* ao, const int64_t* uo, const int64_t* so, const int64_t* vo,
                                      int precision, int* rc) {
    std::vector<int64_t> key;
    key.reserve(2 + 6 * (size_t)count);
    key.push_back(count);
    key.push_back(precision);
    for (int b = 0; b < count; ++b) {
        key.push_back(m[b]);
        key.push_back(n[b]);
        key.push_back(ao[b]);
        key.push_back(uo[b]);
        key.push_back(so[b]);
        key.push_back(vo[b]);
    }
    for (auto& e : h->device_plans)
        if (e.first == key) {
            *rc = LSVD_OK;
            return e.second;
        }
    std::vector<int32_t> modes(count, kModeSvd);
    const int saved = h->ptr_mode;
    h->ptr_mode = LSVD_PTR_DEVICE;
    lsvd_plan_t p = nullptr;
    *rc = plan_build(h, count, m, n, ao, uo, so, vo, modes.data(), precision, &p);
    h->ptr_mode = saved;
    if (*rc != LSVD_OK) return nullptr;
    if (h->device_plans.size() >= 8) {
        lsvd_batch_plan_destroy(h->device_plans.front().second);
        h->device_plans.erase(h->device_plans.begin());
    }
    h->device_plans.emplace_back(std::move(key), p);
    return p;
}

int svd_device_small(lsvd_handle_t h, int count, const int32_t* n, const int64_t* mat_off, const int64_t* s_off,
                     int precision, const double* a, double* u, double* s, double* vt, lsvd_info_t* info) {
    int rc = LSVD_OK;
    lsvd_plan_t p = cached_device_plan(h, count, n, n, mat_off, mat_off, s_off, mat_off, precision, &rc);
    if (!p) return rc;
    rc = plan_execute(p, a, u, s, vt, info, plan_lanes(p) <= 1);
    if (rc == LSVD_OK) {
        hipError_t e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) rc = hip_fail(h, e, "hipStreamSynchronize");
    }
    return rc;
}

int svd_device_uniform(lsvd_handle_t h, int count, int m, int n, const double* a, double* u, double* s, double* vt) {
    std::vector<int32_t> mm(count, m), nn(count, n);
    std::vector<int64_t> ao(count), uo(count), so(count), vo(count);
    const int k = std::min(m, n);
    for (int b = 0; b < count; ++b) {
        ao[b] = (int64_t)b * m * n;
        uo[b] = (int64_t)b * m * k;
        so[b] = (int64_t)b * k;
        vo[b] = (int64_t)b * k * n;
    }
    int rc = LSVD_OK;
    lsvd_plan_t p = cached_device_plan(h, count, mm.data(), nn.data(), ao.data(), uo.data(), so.data(), vo.data(),
                                       LSVD_PREC_F64, &rc);
    if (!p) return rc;
    rc = plan_execute(p, a, u, s, vt, nullptr, false);
    if (rc == LSVD_OK) {
        hipError_t e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) rc = hip_fail(h, e, "hipStreamSynchronize");
    }
    return rc;
}
}  // namespace lsvd

// ---- C ABI ------------------------------------------------------------------

extern "C" {

const char* lsvd_version(void) { return LSVD_VERSION; }

const char* lsvd_status_string(int status) {
    switch (status) {
        case LSVD_OK: return "ok";
        case LSVD_E_SHAPE: return "shape error";
        case LSVD_E_HIP: return "HIP runtime error";
        case LSVD_E_NONFINITE: return "non-finite input";
        case LSVD_E_NOCONV: return "no convergence";
        case LSVD_E_ARG: return "invalid argument";
        case LSVD_E_NOMEM: return "out of device memory";
        case LSVD_E_LIMIT: return "size limit exceeded";
    }
    return "unknown status";
}

int lsvd_create(int device, lsvd_handle_t* out) {
    if (!out) return LSVD_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return LSVD_E_HIP;
    if (device < 0 || device >= ndev) return LSVD_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return LSVD_E_HIP;
    auto* h = new lsvd_handle_s();
    h->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        delete h;
        return LSVD_E_HIP;
    }
    h->num_cus = prop.multiProcessorCount;
    h->arch = prop.gcnArchName;
    if (hipStreamCreateWithFlags(&h->own_stream, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return LSVD_E_HIP;
    }
    h->stream = h->own_stream;
    if (noalloc_init(device) != hipSuccess) {
        (void)hipStreamDestroy(h->own_stream);
        delete h;
        return LSVD_E_HIP;
    }
    h->large_min_n = kLargeMinNDefault;
    if (const char* env = std::getenv("LSVD_LARGE_MIN_N")) {
        char* end = nullptr;
        const long v = std::strtol(env, &end, 10);
        if (end != env && v >= 2) h->large_min_n = (int)std::min<long>(v, kLargeMaxN + 1);
    }
    if (const char* env = std::getenv("LSVD_OS_MAX_NP")) {
        char* end = nullptr;
        const long v = std::strtol(env, &end, 10);
        if (end != env && v >= 0) h->os_max_np = (int)std::min<long>(v, 128);
    }
    if (const char* env = std::getenv("LSVD_BATCH_MULTI")) h->multi_launch = std::strtol(env, nullptr, 10) != 0;
    if (const char* env = std::getenv("LSVD_JADE_CHUNK_MB"))
        h->jade_chunk_bytes = (size_t)std::max(0L, std::strtol(env, nullptr, 10)) << 20;
    if (const char* env = std::getenv("LSVD_LANE0_CUS")) h->lane0_cus = (int)std::strtol(env, nullptr, 10);
    if (const char* env = std::getenv("LSVD_LEAD_NP")) h->lead_np = (int)std::strtol(env, nullptr, 10);
    if (const char* env = std::getenv("LSVD_MICRO_MAX_SWEEPS")) {
        char* end = nullptr;
        const long v = std::strtol(env, &end, 10);
        if (end != env && v >= 1) h->micro_max_sweeps = (int)std::min<long>(v, 24);
    }
    *out = h;
    return LSVD_OK;
}

int lsvd_destroy(lsvd_handle_t h) {
    if (!h) return LSVD_E_ARG;
    (void)hipSetDevice(h->device);
    for (auto& kv : h->plan_cache) lsvd_batch_plan_destroy(kv.second);
    h->plan_cache.clear();
    for (auto& e : h->device_plans) lsvd_batch_plan_destroy(e.second);
    h->device_plans.clear();
    lsvd_release_extras(h);
    delete h->ws;
    h->ws = nullptr;
    if (h->own_stream) (void)hipStreamDestroy(h->own_stream);
    for (hipStream_t s2 : h->lanes) (void)hipStreamDestroy(s2);
    delete h;
    return LSVD_OK;
}

int lsvd_set_stream(lsvd_handle_t h, void* stream) {
    if (!h) return LSVD_E_ARG;
    // the workspace's bytes are reused in stream order: drain the old stream
    const hipStream_t next = stream ? (hipStream_t)stream : h->own_stream;
    if (next != h->stream && h->stream) {
        hipError_t e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) return hip_fail(h, e, "hipStreamSynchronize");
    }
    h->stream = next;
    return LSVD_OK;
}

void* lsvd_get_stream(lsvd_handle_t h) { return h ? (void*)h->stream : nullptr; }

int lsvd_set_pointer_mode(lsvd_handle_t h, int mode) {
    if (!h) return LSVD_E_ARG;
    if (mode != LSVD_PTR_HOST && mode != LSVD_PTR_DEVICE) return set_error(h, LSVD_E_ARG, "unknown pointer mode");
    if (mode != h->ptr_mode) {
        for (auto& kv : h->plan_cache) lsvd_batch_plan_destroy(kv.second);
        h->plan_cache.clear();
    }
    h->ptr_mode = mode;
    return LSVD_OK;
}

int lsvd_set_solve_precision(lsvd_handle_t h, int precision) {
    if (!h) return LSVD_E_ARG;
    if (precision != LSVD_PREC_F64 && precision != LSVD_PREC_F64_SCALED && precision != LSVD_PREC_F64_ONESIDED)
        return set_error(h, LSVD_E_ARG, "single solves take LSVD_PREC_F64, _F64_SCALED or _F64_ONESIDED");
    h->solve_precision = precision;
    return LSVD_OK;
}

int lsvd_synchronize(lsvd_handle_t h) {
    if (!h) return LSVD_E_ARG;
    hipError_t e = hipStreamSynchronize(h->stream);
    return e == hipSuccess ? LSVD_OK : hip_fail(h, e, "hipStreamSynchronize");
}

const char* lsvd_last_error(lsvd_handle_t h) { return h ? h->last_error.c_str() : ""; }

int lsvd_batch_plan_create(lsvd_handle_t h, int count, const int32_t* m, const int32_t* n, const int64_t* a_off,
                           const int64_t* u_off, const int64_t* s_off, const int64_t* vt_off, int precision,
                           lsvd_plan_t* out) {
    if (!h || !out || (count > 0 && (!m || !n || !a_off || !u_off || !s_off || !vt_off))) return LSVD_E_ARG;
    std::vector<int32_t> modes(count);
    for (int b = 0; b < count; ++b) modes[b] = (m[b] == n[b] && n[b] <= 4) ? kModeMicro : kModeSvd;
    return plan_build(h, count, m, n, a_off, u_off, s_off, vt_off, modes.data(), precision, out);
}

int lsvd_batch_plan_execute(lsvd_plan_t p, const double* a_pool, double* u_pool, double* s_pool, double* vt_pool,
                            lsvd_info_t* info) {
    // a plan with concurrent lanes is enqueued on the handle's lanes (graph
    // replays map their parallel branches onto hardware queues of the
    // runtime's choosing, measured 12.4 or 14.6 ms for the same 12288-matrix
    // batch from one process to the next); a one-lane plan replays its graph
    return plan_execute(p, a_pool, u_pool, s_pool, vt_pool, info, plan_lanes(p) <= 1);
}

int lsvd_batch_plan_destroy(lsvd_plan_t p) {
    if (!p) return LSVD_E_ARG;
    (void)hipSetDevice(p->h->device);
    if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    for (hipEvent_t ev : p->ev) (void)hipEventDestroy(ev);
    void* ptrs[] = {p->d_items, p->d_counters, p->d_consts, p->d_gws, p->d_a, p->d_u, p->d_s, p->d_vt, p->d_info};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    delete p;
    return LSVD_OK;
}

int lsvd_svd_batched_f64(lsvd_handle_t h, int count, const int32_t* m, const int32_t* n, const int64_t* a_off,
                         const double* a_pool, const int64_t* u_off, double* u_pool, const int64_t* s_off,
                         double* s_pool, const int64_t* vt_off, double* vt_pool, int precision,
                         lsvd_info_t* info) {
    lsvd_plan_t p = nullptr;
    int rc = lsvd_batch_plan_create(h, count, m, n, a_off, u_off, s_off, vt_off, precision, &p);
    if (rc != LSVD_OK) return rc;
    rc = plan_execute(p, a_pool, u_pool, s_pool, vt_pool, info, false);
    if (rc == LSVD_OK && h->ptr_mode == LSVD_PTR_DEVICE) {
        hipError_t e = hipStreamSynchronize(h->stream);  // the plan's buffers are freed below
        if (e != hipSuccess) rc = hip_fail(h, e, "hipStreamSynchronize");
    }
    lsvd_batch_plan_destroy(p);
    return rc;
}

// The plan-level knobs of the on-chip buckets: a coordinate search over
// the multi-size launch's first padded size, the lane count and the threads
// per workgroup of each per-size bucket, timing the whole plan on `pools`
// (median of 5 executions after a warm-up; a candidate replaces the best so
// far when it saves more than 1.5 %).  Leaves the chosen knobs in the handle
// and the report in h->tune_report.
static int tune_plan_knobs(lsvd_handle_t h, int count, const int32_t* m, const int32_t* n, const int64_t* a_off,
                           const int64_t* u_off, const int64_t* s_off, const int64_t* vt_off, int precision,
                           const Pools& pools, hipEvent_t e0, hipEvent_t e1) {
    const int saved_ptr_mode = h->ptr_mode;
    std::vector<int32_t> pmodes(count);
    for (int b = 0; b < count; ++b) pmodes[b] = (m[b] == n[b] && n[b] <= 4) ? kModeMicro : kModeSvd;
    const int saved_mm = h->multi_min_np, saved_lanes = h->max_lanes, saved_os = h->os_max_np, saved_lead = h->lead_np;
    const uint32_t saved_wide = h->wide_mask;
    int evals = 0;
    float t_ms = 0.f;
    auto time_plan = [&](float* out) -> int {
        h->ptr_mode = LSVD_PTR_DEVICE;
        lsvd_plan_t q = nullptr;
        int r = plan_build(h, count, m, n, a_off, u_off, s_off, vt_off, pmodes.data(), precision, &q);
        h->ptr_mode = saved_ptr_mode;
        if (r != LSVD_OK) return r;
        float ts[5];
        const char* what = "warm-up run";
        hipError_t e = plan_enqueue(q, pools, h->stream);
        for (int rep = 0; rep < 5 && e == hipSuccess; ++rep) {
            what = "event record";
            e = hipEventRecord(e0, h->stream);
            if (e == hipSuccess) {
                what = "timed run";
                e = plan_enqueue(q, pools, h->stream);
            }
            if (e == hipSuccess) {
                what = "event record";
                e = hipEventRecord(e1, h->stream);
            }
            if (e == hipSuccess) {
                what = "event wait";
                e = hipEventSynchronize(e1);
            }
            if (e == hipSuccess) {
                what = "event time";
                e = hipEventElapsedTime(&ts[rep], e0, e1);
            }
        }
        lsvd_batch_plan_destroy(q);
        if (e != hipSuccess) {
            const std::string msg = std::string("autotune plan (multi from ") + std::to_string(h->multi_min_np) +
                                    ", lanes " + std::to_string(h->max_lanes) + ", wide " +
                                    std::to_string(h->wide_mask) + "): " + what;
            return hip_fail(h, e, msg.c_str());
        }
        std::sort(ts, ts + 5);
        *out = ts[2];
        ++evals;
        return LSVD_OK;
    };
    int rc2 = time_plan(&t_ms);
    if (rc2 != LSVD_OK) {
        return rc2;
    }
    const float t_default = t_ms;
    float t_best = t_ms;
    // coordinate search: a candidate replaces the best so far when it
    // saves more than 1.5 %
    int best_mm = saved_mm, best_lanes = saved_lanes, best_os = saved_os, best_lead = saved_lead;
    uint32_t best_wide = saved_wide;
    auto try_knobs = [&](int mm, int lanes, uint32_t wide, int osn, int lead) -> int {
        if (mm == best_mm && lanes == best_lanes && wide == best_wide && osn == best_os && lead == best_lead)
            return LSVD_OK;
        h->multi_min_np = mm;
        h->max_lanes = lanes;
        h->wide_mask = wide;
        h->os_max_np = osn;
        h->lead_np = lead;
        float t = 0.f;
        const int r = time_plan(&t);
        if (r == LSVD_OK && t < t_best * 0.985f) {
            t_best = t;
            best_mm = mm;
            best_lanes = lanes;
            best_wide = wide;
            best_os = osn;
            best_lead = lead;
        }
        h->multi_min_np = best_mm;
        h->max_lanes = best_lanes;
        h->wide_mask = best_wide;
        h->os_max_np = best_os;
        h->lead_np = best_lead;
        return r;
    };
    if (precision == LSVD_PREC_F64_ONESIDED) {  // where QR + one-sided Jacobi hands over to Newton-Schulz
        for (int osn = 16; osn <= 128; osn += 16) {
            const int r = try_knobs(best_mm, best_lanes, best_wide, osn, best_lead);
            if (r != LSVD_OK) return r;
        }
    }
    const int mms[] = {64, 80, 96, 129};
    for (int mm : mms) {
        const int r = try_knobs(mm, best_lanes, best_wide, best_os, best_lead);
        if (r != LSVD_OK) return r;
    }
    for (int lanes = 1; lanes <= kPlanLanes; ++lanes) {
        const int r = try_knobs(best_mm, lanes, best_wide, best_os, best_lead);
        if (r != LSVD_OK) return r;
    }
    for (int np = 16; np <= 64; np += 16) {
        if (np >= best_mm && h->multi_launch) continue;  // inside the multi-size launch
        const int r = try_knobs(best_mm, best_lanes, best_wide ^ (1u << (np / 16 - 1)), best_os, best_lead);
        if (r != LSVD_OK) return r;
    }
    {  // which per-size bucket, if any, goes first
        h->ptr_mode = LSVD_PTR_DEVICE;
        lsvd_plan_t q = nullptr;
        int r = plan_build(h, count, m, n, a_off, u_off, s_off, vt_off, pmodes.data(), precision, &q);
        h->ptr_mode = saved_ptr_mode;
        if (r != LSVD_OK) return r;
        std::vector<int> nps;
        for (const Bucket& bk : q->buckets)
            if (bucket_np(bk) > 0 && bucket_np(bk) != best_lead) nps.push_back(bucket_np(bk));
        lsvd_batch_plan_destroy(q);
        if (best_lead > 0) nps.push_back(0);
        for (int lead : nps) {
            r = try_knobs(best_mm, best_lanes, best_wide, best_os, lead);
            if (r != LSVD_OK) return r;
        }
    }
    h->tune_report = lsvd_tune_report_t{best_mm, best_lanes, (int32_t)best_wide, evals, (double)t_default,
                                        (double)t_best, best_os, best_lead};
        return LSVD_OK;
}

// Autotuner.  Buckets above 128: every compatible kernel variant is timed on
// the bucket alone (best of 3) and the fastest recorded in the handle's
// cache, keyed by (maxN, arena, precision).  Then the plan-level knobs of
// the on-chip buckets (the multi-size launch's first padded size, the lane
// count, default or doubled threads per workgroup for each per-size bucket)
// are chosen by a coordinate search that times the whole plan (median of 5
// executions; a change is kept when it saves more than 1.5 %), since the
// buckets share the CUs and the best choice for one depends on the others.
int lsvd_autotune_batch(lsvd_handle_t h, int count, const int32_t* m, const int32_t* n, int precision) {
    if (!h || count < 0 || (count > 0 && (!m || !n))) return LSVD_E_ARG;
    std::vector<int64_t> a_off(count), u_off(count), s_off(count), vt_off(count);
    int64_t a = 0, u = 0, sv = 0, vt = 0;
    for (int b = 0; b < count; ++b) {
        const int k = std::min(m[b], n[b]);
        a_off[b] = a; u_off[b] = u; s_off[b] = sv; vt_off[b] = vt;
        a += (int64_t)m[b] * n[b]; u += (int64_t)m[b] * k; sv += k; vt += (int64_t)k * n[b];
    }
    const int saved_mode = h->ptr_mode;
    h->ptr_mode = LSVD_PTR_DEVICE;
    lsvd_plan_t p = nullptr;
    std::vector<int32_t> modes(count, kModeSvd);
    int rc = plan_build(h, count, m, n, a_off.data(), u_off.data(), s_off.data(), vt_off.data(), modes.data(),
                        precision, &p);
    h->ptr_mode = saved_mode;
    if (rc != LSVD_OK) return rc;
    std::vector<double> host(std::max<int64_t>(a, 1));
    uint64_t x = 0x5EED7E57ull;
    for (auto& v : host) {  // uniform(-1, 1) from a SplitMix64 stream
        x += 0x9E3779B97F4A7C15ull;
        uint64_t z = x;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        v = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    }
    double *da = nullptr, *du = nullptr, *ds = nullptr, *dvt = nullptr, *gws = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&]() {
        for (double* q : {da, du, ds, dvt, gws})
            if (q) (void)hipFree(q);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        lsvd_batch_plan_destroy(p);
    };
#define TCHECK(expr)                                  \
    do {                                              \
        hipError_t e_ = (expr);                       \
        if (e_ != hipSuccess) {                       \
            int r_ = hip_fail(h, e_, #expr);          \
            cleanup();                                \
            return r_;                                \
        }                                             \
    } while (0)
    TCHECK(hipMalloc(&da, sizeof(double) * host.size()));
    TCHECK(hipMalloc(&du, sizeof(double) * std::max<int64_t>(u, 1)));
    TCHECK(hipMalloc(&ds, sizeof(double) * std::max<int64_t>(sv, 1)));
    TCHECK(hipMalloc(&dvt, sizeof(double) * std::max<int64_t>(vt, 1)));
    TCHECK(hipMemcpy(da, host.data(), sizeof(double) * host.size(), hipMemcpyHostToDevice));
    TCHECK(hipEventCreate(&e0));
    TCHECK(hipEventCreate(&e1));
    Pools pools{da, du, ds, dvt, nullptr};
    int nv = 0;
    const SmallVariant* vs = small_variants(&nv);
    for (size_t bi = 0; bi < p->buckets.size(); ++bi) {
        Bucket base = p->buckets[bi];
        if (base.variant >= kV2Base) continue;  // one on-chip variant per size class
        int best = base.variant;
        float best_ms = 1e30f;
        for (int vi = 0; vi < nv; ++vi) {
            if (!variant_accepts(vs[vi].id, base.maxN, base.arena_bytes)) continue;
            Bucket bk = base;
            const size_t need = configure_bucket(h, bk, vs[vi].id);
            if (gws) {
                (void)hipFree(gws);
                gws = nullptr;
            }
            if (need) TCHECK(hipMalloc(&gws, sizeof(double) * need));
            float t_min = 1e30f;
            for (int rep = 0; rep < 3; ++rep) {
                TCHECK(hipEventRecord(e0, h->stream));
                TCHECK(launch_reset_counters(p->d_counters, p->n_counters, h->stream));
                SmallLaunch L{};
                L.grid = bk.grid;
                L.lds_bytes = bk.lds_bytes;
                L.items = p->d_items + bk.first;
                L.count_ptr = p->d_consts + bk.consts_off;
                L.counter = p->d_counters + bi;
                L.pools = pools;
                L.gws = gws;
                L.gws_stride = bk.gws_stride;
                L.ns_scaled = precision == LSVD_PREC_F64_SCALED;
                TCHECK(launch_small_variant(bk.variant, L, h->stream));
                TCHECK(hipEventRecord(e1, h->stream));
                TCHECK(hipEventSynchronize(e1));
                float ms = 0.f;
                TCHECK(hipEventElapsedTime(&ms, e0, e1));
                t_min = std::min(t_min, ms);
            }
            if (t_min < best_ms) {
                best_ms = t_min;
                best = vs[vi].id;
            }
        }
        h->tune[tune_key(base.maxN, base.arena_bytes, precision)] = best;
    }
    // ---- plan-level knobs, timing whole plans on the same inputs
    rc = tune_plan_knobs(h, count, m, n, a_off.data(), u_off.data(), s_off.data(), vt_off.data(), precision, pools,
                         e0, e1);
    if (rc != LSVD_OK) {
        cleanup();
        return rc;
    }
#undef TCHECK
    cleanup();
    return LSVD_OK;
}

int lsvd_autotune_batch_pools(lsvd_handle_t h, int count, const int32_t* m, const int32_t* n, const int64_t* a_off,
                              const int64_t* u_off, const int64_t* s_off, const int64_t* vt_off, int precision,
                              const double* a_pool, double* u_pool, double* s_pool, double* vt_pool) {
    if (!h || count < 0 || (count > 0 && (!m || !n || !a_off || !u_off || !s_off || !vt_off))) return LSVD_E_ARG;
    if (h->ptr_mode != LSVD_PTR_DEVICE) return set_error(h, LSVD_E_ARG, "lsvd_autotune_batch_pools takes device pools");
    if (count > 0 && (!a_pool || !u_pool || !s_pool || !vt_pool)) return set_error(h, LSVD_E_ARG, "null pool pointer");
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(h, e, "hipSetDevice");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if ((e = hipEventCreate(&e0)) != hipSuccess) return hip_fail(h, e, "hipEventCreate");
    if ((e = hipEventCreate(&e1)) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return hip_fail(h, e, "hipEventCreate");
    }
    const Pools pools{a_pool, u_pool, s_pool, vt_pool, nullptr};
    const int rc = tune_plan_knobs(h, count, m, n, a_off, u_off, s_off, vt_off, precision, pools, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int lsvd_autotune_report(lsvd_handle_t h, lsvd_tune_report_t* out) {
    if (!h || !out) return LSVD_E_ARG;
    *out = h->tune_report;
    return LSVD_OK;
}

int lsvd_batch_plan_describe(lsvd_plan_t p, lsvd_bucket_desc_t* out, int max, int* nbuckets, int32_t* order) {
    if (!p || !nbuckets) return LSVD_E_ARG;
    std::vector<lsvd_bucket_desc_t> d;
    if (p->n_micro > 0) d.push_back({0, p->n_micro, 4, -1, (p->n_micro + 255) / 256, 256, 0});
    for (auto& bk : p->buckets) {
        const SmallVariant* v = bk.variant >= kV2Base ? nullptr : find_variant(bk.variant);
        const int nt = bk.variant >= kOsMulti   ? 512
                       : bk.variant >= kOsBase  ? os_threads(bk.variant - kOsBase)
                       : bk.variant >= kV2Wide  ? 2 * v2_threads(bk.variant - kV2Wide)
                       : bk.variant >= kV2Multi ? 1024
                       : bk.variant >= kV2Base  ? v2_threads(bk.variant - kV2Base)
                                                : (v ? v->nt : 0);
        d.push_back({bk.first, bk.count, bk.maxN, bk.variant, bk.grid, nt, (int64_t)bk.lds_bytes});
    }
    *nbuckets = (int)d.size();
    for (int i = 0; i < (int)d.size() && i < max && out; ++i) out[i] = d[i];
    if (order) {
        std::vector<Item> items(p->fallback_first);
        if (p->fallback_first) {
            hipError_t e = hipMemcpy(items.data(), p->d_items, sizeof(Item) * items.size(), hipMemcpyDeviceToHost);
            if (e != hipSuccess) return hip_fail(p->h, e, "hipMemcpy");
        }
        for (int i = 0; i < p->fallback_first; ++i) order[i] = items[i].idx;
    }
    return LSVD_OK;
}

int lsvd_batch_plan_time_buckets(lsvd_plan_t p, const double* a_pool, double* u_pool, double* s_pool,
                                 double* vt_pool, float* ms, int max) {
    if (!p || !ms) return LSVD_E_ARG;
    lsvd_handle_t h = p->h;
    if (p->ptr_mode != LSVD_PTR_DEVICE) return set_error(h, LSVD_E_ARG, "time_buckets needs device pointers");
    hipStream_t st = h->stream;
    Pools pools{a_pool, u_pool, s_pool, vt_pool, nullptr};
    const int nb = (int)p->buckets.size();
    // one run of the execute schedule with start/stop events around every
    // launch: each duration is measured while the other lanes run beside it
    std::vector<hipEvent_t> ev(2 * (nb + 1), nullptr);
    int rc = LSVD_OK;
    for (auto& x : ev)
        if (hipEventCreate(&x) != hipSuccess) rc = set_error(h, LSVD_E_HIP, "hipEventCreate");
    if (rc == LSVD_OK) {
        hipError_t e = plan_enqueue(p, pools, st, ev.data());
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = hip_fail(h, e, "bucket timing");
    }
    int slot = 0;
    for (int i = (p->n_micro > 0 ? 0 : 1); i <= nb && rc == LSVD_OK; ++i) {
        float t = 0.f;
        hipError_t e = hipEventElapsedTime(&t, ev[2 * i], ev[2 * i + 1]);
        if (e != hipSuccess) rc = hip_fail(h, e, "hipEventElapsedTime");
        if (slot < max) ms[slot] = t;
        ++slot;
    }
    for (auto& x : ev)
        if (x) (void)hipEventDestroy(x);
    return rc;
}

int lsvd_batch_plan_phase_clocks(lsvd_plan_t p, const double* a_pool, double* u_pool, double* s_pool,
                                 double* vt_pool, int64_t* clocks) {
    if (!p || !clocks) return LSVD_E_ARG;
    lsvd_handle_t h = p->h;
    if (p->ptr_mode != LSVD_PTR_DEVICE) return set_error(h, LSVD_E_ARG, "phase_clocks needs device pointers");
    Pools pools{a_pool, u_pool, s_pool, vt_pool, nullptr, reinterpret_cast<long long*>(clocks)};
    hipError_t e = plan_enqueue(p, pools, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    return e == hipSuccess ? LSVD_OK : hip_fail(h, e, "phase clocks run");
}


// Single-matrix entry points: a cached one-item plan per (mode, m, n).
static int single(lsvd_handle_t h, int mode, const double* a, int m, int n, double* u, double* s, double* vt,
                  lsvd_info_t* info) {
    if (!h) return LSVD_E_ARG;
    if (m < 0 || n < 0) return set_error(h, LSVD_E_SHAPE, "negative dimension");
    const int k = std::min(m, n);
    if (mode == kModeSvd && (m == 0 || n == 0))
        return set_error(h, LSVD_E_SHAPE, "LieSvdSmall::solve_rectangular: empty input");
    if (k == 0 && mode != kModeSvd) {
        if (info) *info = lsvd_info_t{};
        return LSVD_OK;
    }
    if (!a || !u || !s || (mode != kModeEigh && !vt)) return set_error(h, LSVD_E_ARG, "null pointer");
    if (((mode == kModeSvd && k >= h->large_min_n) || (mode == kModeEigh && m == n && n >= h->large_min_n)) &&
        k <= kLargeMaxN) {
        hipError_t e = hipSetDevice(h->device);
        if (e != hipSuccess) return hip_fail(h, e, "hipSetDevice");
        bool done = false;
        int rc = mode == kModeSvd ? svd_large(h, a, m, n, u, s, vt, info, &done)
                                  : eigh_large(h, a, n, u, s, info, &done);
        if (rc != LSVD_OK || done) return rc;
    }
    const uint64_t key = ((uint64_t)mode << 60) | ((uint64_t)(uint32_t)h->solve_precision << 54) |
                         ((uint64_t)(uint32_t)m << 27) | (uint64_t)(uint32_t)n;
    lsvd_plan_t p = nullptr;
    auto it = h->plan_cache.find(key);
    if (it != h->plan_cache.end()) {
        p = it->second;
    } else {
        int32_t mm = m, nn = n, md = mode;
        int64_t zero = 0;
        int rc = plan_build(h, 1, &mm, &nn, &zero, &zero, &zero, &zero, &md, h->solve_precision, &p);
        if (rc != LSVD_OK) return rc;
        if (h->plan_cache.size() > 64) {
            for (auto& kv : h->plan_cache) lsvd_batch_plan_destroy(kv.second);
            h->plan_cache.clear();
        }
        h->plan_cache[key] = p;
    }
    lsvd_info_t tmp_info;
    int rc = plan_execute(p, a, u, s, vt, info ? (h->ptr_mode == LSVD_PTR_HOST ? &tmp_info : info) : nullptr, false);
    if (rc == LSVD_OK && info && h->ptr_mode == LSVD_PTR_HOST) *info = tmp_info;
    return rc;
}

int lsvd_small_solve_f64(lsvd_handle_t h, const double* a, int n, double* u, double* s, double* vt,
                         lsvd_info_t* info) {
    if (!h) return LSVD_E_ARG;
    if (n == 0) {  // the reference's solve accepts 0 x 0 and returns empty factors
        if (info) *info = lsvd_info_t{};
        return LSVD_OK;
    }
    return single(h, kModeSvd, a, n, n, u, s, vt, info);
}

int lsvd_small_solve_rectangular_f64(lsvd_handle_t h, const double* a, int m, int n, double* u, double* s,
                                     double* vt, lsvd_info_t* info) {
    return single(h, kModeSvd, a, m, n, u, s, vt, info);
}

int lsvd_eigh_jacobi_f64(lsvd_handle_t h, const double* p, int n, double* v, double* eig, lsvd_info_t* info) {
    return single(h, kModeEigh, p, n, n, v, eig, nullptr, info);
}

int lsvd_micro_solve_f64(lsvd_handle_t h, const double* a, int n, double* u, double* s, double* vt,
                         lsvd_info_t* info) {
    if (h && n > 4) return set_error(h, LSVD_E_SHAPE, "LieSvdMicro only supports N <= 4");
    return single(h, kModeMicro, a, n, n, u, s, vt, info);
}

}  // extern "C"
