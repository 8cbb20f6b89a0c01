// lsvd_kernels.hpp -- host-visible launch descriptors of the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "lsvd_types.hpp"

namespace lsvd {

// ---- batched LieSvdSmall (small_svd.hip) --------------------------------

struct SmallVariant {
    int id;
    int nt;     // threads per workgroup
    int maxn;   // largest N (after transposing wide inputs) it accepts
    bool lds;   // arena in LDS (true) or in a global per-workgroup slot
};

struct SmallLaunch {
    int grid;
    size_t lds_bytes;      // dynamic LDS (arena) per workgroup, LDS variants
    const Item* items;     // device, sorted work queue
    const int* count_ptr;  // device, number of queue entries
    int* counter;          // device, zeroed before the launch
    Pools pools;
    double* gws;           // global arena slots (global variants)
    int64_t gws_stride;    // doubles per slot
    int ns_scaled;         // scaled Newton-Schulz schedule (on-chip kernel)
};

// Arena rows for a tall-ified problem M x N (M >= N); ld = N | 1.
inline int small_arena_rows(int M, int N) { return (3 * N > M + N) ? 3 * N : M + N; }
inline size_t small_arena_doubles(int M, int N) { return (size_t)small_arena_rows(M, N) * (size_t)(N | 1); }

const SmallVariant* small_variants(int* count);
hipError_t launch_small_variant(int id, const SmallLaunch& L, hipStream_t st);
int small_variant_blocks_per_cu(int id, size_t lds_bytes);
hipError_t launch_reset_counters(int* counters, int n, hipStream_t st);

// ---- on-chip batched LieSvdSmall for N, M <= 128 (small_svd2.hip) -------
// np: padded size 16, 32, ..., 128 (N <= np).  Variant id = 100 + np with
// the default threads per workgroup, 500 + np with twice as many (np <= 64).
int v2_threads(int np);
bool v2_has_wide(int np);
size_t v2_lds_bytes(int np, int max_m);
int64_t v2_slot_doubles(int np, int max_m);
hipError_t launch_small2(int np, bool wide, const SmallLaunch& L, hipStream_t st);
int small2_blocks_per_cu(int np, bool wide, size_t lds_bytes);
constexpr int kV2Base = 100;
constexpr int kV2Wide = 500;
// One persistent launch over the items of every padded size 64..128 (1024
// threads per matrix, LDS and slots sized for np_max): variant id =
// kV2Multi + np_max.
constexpr int kV2Multi = 300;
constexpr int kV2MultiMinNP = 64;
hipError_t launch_small2_multi(const SmallLaunch& L, hipStream_t st);
int small2_multi_blocks_per_cu(size_t lds_bytes);

// ---- QR + one-sided Jacobi for N, M <= 128 (onesided.hip) --------------
// np: padded size 16, 32, ..., 128.  Variant id = kOsBase + np.
constexpr int kOsBase = 700;
int os_threads(int np);
size_t os_lds_bytes(int np, int max_m);
int64_t os_slot_doubles(int np, int max_m);
hipError_t launch_onesided(int np, const SmallLaunch& L, hipStream_t st);
int onesided_blocks_per_cu(int np, size_t lds_bytes);
// One persistent launch over the one-sided items of padded sizes 80..128
// (512 threads, LDS and slot sized for the largest): variant id =
// kOsMulti + np_max; count_ptr[1..4] end the classes 128, 112, 96, 80.
constexpr int kOsMulti = 900;
constexpr int kOsMultiMinNP = 80;
hipError_t launch_onesided_multi(const SmallLaunch& L, hipStream_t st);
int onesided_multi_blocks_per_cu(size_t lds_bytes);

// ---- LieSvdMicro (micro_svd.hip) ----------------------------------------

struct MicroLaunch {
    const Item* items;     // device, count entries, all n <= 4 square
    int count;
    Pools pools;
    Item* fallback;        // device, items escalated to LieSvdSmall
    int* fallback_count;   // device, zeroed before the launch
    int sweep_cap = 24;    // sweep limit of n = 3, 4 (the reference's 24)
};

hipError_t launch_micro(const MicroLaunch& L, hipStream_t st);

}  // namespace lsvd
