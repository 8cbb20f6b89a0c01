// lsvd_schedule.hpp -- launch interface of the Givens export kernel
// (givens.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lsvd {

// One matrix of the batch: offsets in doubles (src, into the input pool),
// in event slots (ev) and in sign slots (sg); idx is its position in the
// caller's list (for n_events).
struct GivensItem {
    int64_t src;
    int64_t ev;
    int64_t sg;
    int d;
    int idx;
};

struct GivensLaunch {
    const GivensItem* items;  // device, `count` entries of one size class
    int count;
    const double* pool;       // device input pool
    int transpose;            // eliminate the transpose of each matrix
    int* n_events;            // device, one per caller item
    int* ev_ij;               // device, 2 per event slot
    double* ev_cs;            // device, 2 per event slot
    double* ev_theta;         // device or null
    double* signs;            // device
    double* scratch;          // global working copies (classes above kGivensLdsMaxD)
    int64_t scratch_stride;   // doubles per workgroup slot
};

constexpr int kGivensLdsMaxD = 128;   // working copy in LDS up to here
constexpr int kGivensMaxD = 1024;     // 16 columns per lane

// d_max: largest d in the launch (selects columns per lane and the LDS
// bytes); grid: workgroups (each loops over items grid apart).
hipError_t launch_givens(const GivensLaunch& L, int d_max, int grid, hipStream_t st);

}  // namespace lsvd
