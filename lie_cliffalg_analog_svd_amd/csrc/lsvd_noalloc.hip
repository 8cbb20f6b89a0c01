// lsvd_noalloc.hip -- an allocation-free variant of
// lsvd_small_solve_rectangular_f64 for callers that cannot allocate device
// memory (code under stream capture, serving loops, harnesses that forbid
// hipMalloc).
//
// The solve's work-queue entry, its counter and queue constants, and the
// workgroup's staging slot live in `scratch`, a file-scope device array of
// kNoallocScratchBytes that every GPU the library runs on holds until the
// process ends.  Its address is looked up once per device when a handle is
// created (noalloc_init, called by lsvd_create with the device current) and
// kept per device; host code only ever passes that address, never the
// array's name.  The entry point enqueues two kernels on the handle's stream
// (the queue set-up and the on-chip solve of the same variant the
// allocating entry point picks, so the results are identical) and calls no
// allocating routine.  All calls on one GPU share its array: calls that
// overlap on one GPU (two streams, or two host threads) race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "lsvd_api_internal.hpp"

namespace lsvd {

// Largest problem: max(m, n) <= 128, the on-chip kernels' limit.  Layout:
//   [0, 256)   the queue entry (Item)
//   [256, 512) int counter, then the queue constants: the count, and the
//              end of each size class of the multi-size kernel
//   [512, ...) the workgroup's staging slot, v2_slot_doubles(np, M) <=
//              3 * 128 * 128 doubles
constexpr int kNoallocMaxDim = 128;
constexpr size_t kNoallocHead = 512;
constexpr size_t kNoallocScratchBytes =
    kNoallocHead + sizeof(double) * 3 * kNoallocMaxDim * kNoallocMaxDim;  // 393,728 B

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

constexpr int kNoallocMaxDevices = 64;
static std::mutex g_noalloc_mu;
static void* g_noalloc_ptr[kNoallocMaxDevices];

hipError_t noalloc_init(int device) {
    if (device < 0 || device >= kNoallocMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return hipSuccess;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!ptr) return hipErrorInvalidSymbol;
    g_noalloc_ptr[device] = ptr;
    return hipSuccess;
}

static void* noalloc_ptr(int device) {
    if (device < 0 || device >= kNoallocMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    return g_noalloc_ptr[device];
}

struct NoallocSetup {
    Item item;
    int ints[8];  // counter (0), count, class ends
};
static_assert(sizeof(Item) <= 256, "queue entry exceeds its scratch region");

// Writes the queue entry and the counter/constants with vector stores (a
// kernel rather than a copy or memset, so a captured call records plain
// kernel nodes and the counter starts at zero on every call).
__global__ void noalloc_setup_kernel(NoallocSetup a, Item* item, int* ints) {
    if (threadIdx.x == 0) *item = a.item;
    if (threadIdx.x < 8) ints[threadIdx.x] = a.ints[threadIdx.x];
}

}  // namespace lsvd

using namespace lsvd;

extern "C" int lsvd_small_solve_rectangular_f64_noalloc(lsvd_handle_t h, const double* a, int m, int n, double* u,
                                                        double* s, double* vt, lsvd_info_t* info) {
    if (!h) return LSVD_E_ARG;
    if (h->ptr_mode != LSVD_PTR_DEVICE)
        return set_error(h, LSVD_E_ARG,
                         "lsvd_small_solve_rectangular_f64_noalloc takes device pointers (LSVD_PTR_DEVICE)");
    if (m < 0 || n < 0) return set_error(h, LSVD_E_SHAPE, "negative dimension");
    if (m == 0 || n == 0) return set_error(h, LSVD_E_SHAPE, "LieSvdSmall::solve_rectangular: empty input");
    if (m > kNoallocMaxDim || n > kNoallocMaxDim)
        return set_error(h, LSVD_E_LIMIT,
                         "lsvd_small_solve_rectangular_f64_noalloc: " + std::to_string(m) + " x " + std::to_string(n) +
                             " exceeds its fixed scratch (max(m, n) <= 128)");
    if (!a || !u || !s || !vt) return set_error(h, LSVD_E_ARG, "null pointer");
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(h, e, "hipSetDevice");
    char* base = static_cast<char*>(noalloc_ptr(h->device));
    if (!base) return set_error(h, LSVD_E_HIP, "no-alloc scratch of this device is not initialised");

    const int M = m > n ? m : n, N = m < n ? m : n;
    const int np = (N + 15) / 16 * 16;
    const bool multi = onchip_multi(h, np);  // as the batch plan buckets it
    NoallocSetup su{};
    su.item = Item{m, n, 0, kModeSvd, 0, 0, 0, 0};
    su.ints[0] = 0;  // queue counter
    su.ints[1] = 1;  // queue length
    if (multi)       // multi-size kernel: end of the classes 128, 112, 96, 80, 64
        for (int c = 0, cls = 128; cls >= kV2MultiMinNP; ++c, cls -= 16) su.ints[2 + c] = np >= cls ? 1 : 0;
    Item* d_item = reinterpret_cast<Item*>(base);
    int* d_ints = reinterpret_cast<int*>(base + 256);
    hipLaunchKernelGGL(noalloc_setup_kernel, dim3(1), dim3(64), 0, h->stream, su, d_item, d_ints);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(h, e, "noalloc set-up launch");

    SmallLaunch L{};
    L.grid = 1;
    L.lds_bytes = v2_lds_bytes(np, M);
    L.items = d_item;
    L.counter = d_ints;
    L.count_ptr = d_ints + 1;
    L.pools = Pools{a, u, s, vt, info, nullptr};
    L.gws = reinterpret_cast<double*>(base + kNoallocHead);
    L.gws_stride = v2_slot_doubles(np, M);
    L.ns_scaled = 0;  // LSVD_PREC_F64, as the single-matrix entry points
    if (kNoallocHead + sizeof(double) * (size_t)L.gws_stride > kNoallocScratchBytes)
        return set_error(h, LSVD_E_LIMIT, "no-alloc staging slot exceeds the scratch");
    e = launch_small_variant(multi ? kV2Multi + np : onchip_variant(h, np), L, h->stream);
    return e == hipSuccess ? LSVD_OK : hip_fail(h, e, "kernel launch");
}
