// lsvd_schedule.cpp -- host side of the rotor-stream export: the batched
// Givens elimination entry point, the events it stands for, and the
// phase-schedule-v1 JSON writer (src/lie_svd_compiler.rs).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "lsvd_api_internal.hpp"
#include "lsvd_schedule.hpp"

using namespace lsvd;

namespace {

// temporaries come from the handle's workspace (lsvd_workspace.hpp)
using DevMem = lsvd::WsMem;

// Size classes of one launch each: bands of 16 up to the LDS limit (the
// LDS bytes, hence the occupancy, follow the band's largest d), then the
// global-slot classes by columns per lane.
int size_class(int d) {
    if (d <= kGivensLdsMaxD) return d <= 16 ? 0 : (d - 1) / 16;
    if (d <= 256) return 8;
    if (d <= 512) return 9;
    return 10;
}
constexpr int kClasses = 11;

// ---- Rust number formatting -------------------------------------------

// Rust's {:.17e}: the exact value rounded to 18 significant digits, 'e',
// the exponent with '-' only when negative and no zero padding; -0.0 keeps
// its sign; NaN -> "NaN", infinities -> "inf" / "-inf".
void put_exp17(std::string& out, double x) {
    if (std::isnan(x)) {
        out += "NaN";
        return;
    }
    if (std::isinf(x)) {
        out += x < 0 ? "-inf" : "inf";
        return;
    }
    char t[48];
    auto r = std::to_chars(t, t + sizeof t, x, std::chars_format::scientific, 17);
    char* e = std::find(t, r.ptr, 'e');
    out.append(t, e);
    out += 'e';
    const char* p = e + 1;
    if (*p == '-') out += '-';
    if (*p == '+' || *p == '-') ++p;
    while (p + 1 < r.ptr && *p == '0') ++p;
    out.append(p, (size_t)(r.ptr - p));
}

// Rust's {} for f64: the shortest digits that round-trip, written as a
// plain decimal (no exponent, no trailing ".0"); NaN -> "NaN".
void put_display(std::string& out, double x) {
    if (std::isnan(x)) {
        out += "NaN";
        return;
    }
    if (std::isinf(x)) {
        out += x < 0 ? "-inf" : "inf";
        return;
    }
    char t[48];
    auto r = std::to_chars(t, t + sizeof t, x, std::chars_format::scientific);
    char* e = std::find(t, r.ptr, 'e');
    std::string digits;
    const char* p = t;
    if (*p == '-') {
        out += '-';
        ++p;
    }
    for (; p < e; ++p)
        if (*p != '.') digits += *p;
    const int ex = std::atoi(e + 1);  // value = 0.d1 d2 ... * 10^(ex + 1)
    const int point = ex + 1;        // digits before the decimal point
    if (point <= 0) {
        out += "0.";
        out.append((size_t)(-point), '0');
        out += digits;
    } else if ((size_t)point >= digits.size()) {
        out += digits;
        out.append((size_t)point - digits.size(), '0');
    } else {
        out.append(digits, 0, (size_t)point);
        out += '.';
        out.append(digits, (size_t)point, std::string::npos);
    }
}

const char* target_name(int t) {
    switch (t) {
        case LSVD_TARGET_MZI_MESH: return "mzi_mesh";
        case LSVD_TARGET_FPGA_ROTOR_MESH: return "fpga_rotor_mesh";
        default: return nullptr;
    }
}

const char* source_name(int s) {
    switch (s) {
        case LSVD_SRC_REAL_PHASEFLOW: return "real_phaseflow";
        case LSVD_SRC_COMPLEX_PHASEFLOW: return "complex_phaseflow";
        case LSVD_SRC_ORTHOGONAL_GIVENS: return "orthogonal_givens";
        default: return nullptr;
    }
}

const char* kind_name(int k) {
    static const char* const names[] = {
        "golden_prespin", "causal_antispin", "cross_phase_yinyang", "phase_conjugate",
        "bottleneck", "phase_jump", "unwrap", "directional", "surgery",
        "complex_golden_prespin", "complex_diagonal_phase", "givens_elimination"};
    return (k >= 0 && k < (int)(sizeof names / sizeof names[0])) ? names[k] : nullptr;
}

}  // namespace

extern "C" {

int lsvd_givens_schedule_f64(lsvd_handle_t h, int count, const int32_t* d, const int64_t* v_off, const double* v,
                             int transpose, int32_t* n_events, int32_t* ev_ij, double* ev_cs, double* ev_theta,
                             double* signs) {
    if (!h) return LSVD_E_ARG;
    if (count < 0) return set_error(h, LSVD_E_SHAPE, "negative count");
    if (count == 0) return LSVD_OK;
    if (!d || !v_off || !v || !n_events || !ev_ij || !ev_cs || !signs) return set_error(h, LSVD_E_ARG, "null pointer");
    std::vector<GivensItem> items((size_t)count);
    std::vector<int> cls_count(kClasses, 0), cls_dmax(kClasses, 0);
    int64_t ev_total = 0, sg_total = 0, extent = 0;
    for (int b = 0; b < count; ++b) {
        const int db = d[b];
        if (db < 0 || v_off[b] < 0) return set_error(h, LSVD_E_SHAPE, "negative size or offset");
        if (db > kGivensMaxD) return set_error(h, LSVD_E_LIMIT, "givens schedule: d > 1024");
        items[b] = GivensItem{v_off[b], ev_total, sg_total, db, b};
        ev_total += (int64_t)db * (db - 1) / 2;
        sg_total += db;
        extent = std::max(extent, v_off[b] + (int64_t)db * db);
        const int c = size_class(db);
        cls_count[c]++;
        cls_dmax[c] = std::max(cls_dmax[c], db);
    }
    // group the items by class (stable: matrix order inside a class)
    std::vector<int> first(kClasses + 1, 0);
    for (int c = 0; c < kClasses; ++c) first[c + 1] = first[c] + cls_count[c];
    std::vector<GivensItem> sorted((size_t)count);
    {
        std::vector<int> fill(first.begin(), first.end() - 1);
        for (const auto& it : items) sorted[fill[size_class(it.d)]++] = it;
    }
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return hip_fail(h, e, "hipSetDevice");
    hipStream_t st = h->stream;
#define GCHECK(expr)                                        \
    do {                                                    \
        hipError_t e_ = (expr);                             \
        if (e_ != hipSuccess) return hip_fail(h, e_, #expr); \
    } while (0)
    const bool host = h->ptr_mode == LSVD_PTR_HOST;
    lsvd::WsScope ws_scope(h);
    DevMem bitems, bpool, bn, bij, bcs, bth, bsg, bscratch;
    GCHECK(bitems.alloc(sizeof(GivensItem) * (size_t)count));
    GCHECK(hipMemcpyAsync(bitems.p, sorted.data(), sizeof(GivensItem) * (size_t)count, hipMemcpyHostToDevice, st));
    const double* d_pool = v;
    int32_t* d_n = n_events;
    int32_t* d_ij = ev_ij;
    double* d_cs = ev_cs;
    double* d_th = ev_theta;
    double* d_sg = signs;
    if (host) {
        GCHECK(bpool.alloc(sizeof(double) * (size_t)extent));
        GCHECK(hipMemcpyAsync(bpool.p, v, sizeof(double) * (size_t)extent, hipMemcpyHostToDevice, st));
        d_pool = (const double*)bpool.p;
        GCHECK(bn.alloc(sizeof(int32_t) * (size_t)count));
        GCHECK(bij.alloc(sizeof(int32_t) * 2 * (size_t)ev_total));
        GCHECK(bcs.alloc(sizeof(double) * 2 * (size_t)ev_total));
        if (ev_theta) GCHECK(bth.alloc(sizeof(double) * (size_t)ev_total));
        GCHECK(bsg.alloc(sizeof(double) * (size_t)sg_total));
        d_n = (int32_t*)bn.p;
        d_ij = (int32_t*)bij.p;
        d_cs = (double*)bcs.p;
        d_th = ev_theta ? (double*)bth.p : nullptr;
        d_sg = (double*)bsg.p;
    }
    // global working copies for the classes above the LDS limit
    const int glob_grid_cap = 2 * std::max(h->num_cus, 1);
    int64_t scratch_doubles = 0;
    for (int c = 8; c < kClasses; ++c)
        if (cls_count[c])
            scratch_doubles = std::max(scratch_doubles, (int64_t)std::min(cls_count[c], glob_grid_cap) *
                                                            cls_dmax[c] * cls_dmax[c]);
    if (scratch_doubles) GCHECK(bscratch.alloc(sizeof(double) * (size_t)scratch_doubles));
    for (int c = 0; c < kClasses; ++c) {
        if (!cls_count[c]) continue;
        GivensLaunch L{};
        L.items = (const GivensItem*)bitems.p + first[c];
        L.count = cls_count[c];
        L.pool = d_pool;
        L.transpose = transpose ? 1 : 0;
        L.n_events = d_n;
        L.ev_ij = d_ij;
        L.ev_cs = d_cs;
        L.ev_theta = d_th;
        L.signs = d_sg;
        const int dm = std::max(cls_dmax[c], 1);
        int grid = cls_count[c];
        if (dm > kGivensLdsMaxD) {
            grid = std::min(grid, glob_grid_cap);
            L.scratch = (double*)bscratch.p;
            L.scratch_stride = (int64_t)dm * dm;
        }
        GCHECK(launch_givens(L, dm, grid, st));
    }
    if (host) {
        GCHECK(hipMemcpyAsync(n_events, d_n, sizeof(int32_t) * (size_t)count, hipMemcpyDeviceToHost, st));
        if (ev_total) {
            GCHECK(hipMemcpyAsync(ev_ij, d_ij, sizeof(int32_t) * 2 * (size_t)ev_total, hipMemcpyDeviceToHost, st));
            GCHECK(hipMemcpyAsync(ev_cs, d_cs, sizeof(double) * 2 * (size_t)ev_total, hipMemcpyDeviceToHost, st));
            if (ev_theta)
                GCHECK(hipMemcpyAsync(ev_theta, d_th, sizeof(double) * (size_t)ev_total, hipMemcpyDeviceToHost, st));
        }
        if (sg_total) GCHECK(hipMemcpyAsync(signs, d_sg, sizeof(double) * (size_t)sg_total, hipMemcpyDeviceToHost, st));
        GCHECK(hipStreamSynchronize(st));
    } else {
        GCHECK(hipStreamSynchronize(st));  // the temporaries are freed on return
    }
#undef GCHECK
    return LSVD_OK;
}

int lsvd_givens_events(int n, const int32_t* ev_ij, const double* ev_cs, lsvd_phase_event_t* out) {
    if (n < 0) return LSVD_E_SHAPE;
    if (n && (!ev_ij || !ev_cs || !out)) return LSVD_E_ARG;
    for (int e = 0; e < n; ++e) {
        lsvd_phase_event_t& p = out[e];
        const double theta = std::atan2(ev_cs[2 * e + 1], ev_cs[2 * e]);
        p.layer = e;
        p.i = ev_ij[2 * e];
        p.j = ev_ij[2 * e + 1];
        p.flags = 0;
        p.phi_l = 0.0;
        p.phi_r = 0.0;
        p.theta = theta;
        p.theta_l = theta;
        p.theta_r = 0.0;
        p.energy_before = 0.0;
        p.energy_after = 0.0;
        p.source = LSVD_SRC_ORTHOGONAL_GIVENS;
        p.kind = LSVD_KIND_GIVENS_ELIMINATION;
    }
    return LSVD_OK;
}

size_t lsvd_schedule_json(int target, int channels, int n_events, const lsvd_phase_event_t* events, int n_signs,
                          const double* signs, char* buf, size_t cap) {
    const char* tname = target_name(target);
    if (!tname || n_events < 0 || n_signs < 0 || (n_events && !events) || (n_signs && !signs)) return (size_t)-1;
    std::string out;
    out.reserve(160 + (size_t)n_events * 260 + (size_t)n_signs * 4);
    out += "{\n  \"format_version\": \"phase-schedule-v1\",\n  \"target\": \"";
    out += tname;
    out += "\",\n  \"channels\": ";
    out += std::to_string(channels);
    out += ",\n  \"total_events\": ";
    out += std::to_string(n_events);
    out += ",\n  \"events\": [\n";
    for (int e = 0; e < n_events; ++e) {
        const lsvd_phase_event_t& p = events[e];
        const char* src = source_name(p.source);
        const char* kind = kind_name(p.kind);
        if (!src || !kind) return (size_t)-1;
        out += "    {\"layer\": ";
        out += std::to_string(p.layer);
        out += ", \"i\": ";
        out += std::to_string(p.i);
        out += ", \"j\": ";
        out += std::to_string(p.j);
        out += ", \"phi_l\": ";
        put_exp17(out, p.phi_l);
        out += ", \"phi_r\": ";
        put_exp17(out, p.phi_r);
        out += ", \"theta\": ";
        put_exp17(out, p.theta);
        out += ", \"theta_l\": ";
        put_exp17(out, p.theta_l);
        out += ", \"theta_r\": ";
        put_exp17(out, p.theta_r);
        out += ", \"source\": \"";
        out += src;
        out += "\", \"kind\": \"";
        out += kind;
        out += '"';
        if (p.flags & 1) {
            out += ", \"energy_before\": ";
            put_exp17(out, p.energy_before);
        }
        if (p.flags & 2) {
            out += ", \"energy_after\": ";
            put_exp17(out, p.energy_after);
        }
        out += '}';
        if (e + 1 != n_events) out += ',';
        out += '\n';
    }
    out += "  ]";
    if (n_signs > 0) {
        out += ",\n  \"diagonal_signs\": [";
        for (int k = 0; k < n_signs; ++k) {
            if (k) out += ", ";
            put_display(out, signs[k]);
        }
        out += ']';
    }
    out += "\n}\n";
    if (buf && cap) {
        const size_t n = std::min(out.size(), cap - 1);
        std::memcpy(buf, out.data(), n);
        buf[n] = 0;
    }
    return out.size();
}

}  // extern "C"
