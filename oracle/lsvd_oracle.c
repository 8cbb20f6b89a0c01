/*
 * lsvd_oracle.c -- CPU FP64 restatement of the reference dense SVD path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker or as the timed CPU baseline.  The product library
 * (lie_cliffalg_analog_svd_amd) never links, loads or calls it.
 *
 * The reference is the Rust crate skovnats/lie_cliffalg_analog_svd v0.44.0.
 * Every function cites the file:line it restates.  Loop orders, constants,
 * stopping tests, fallbacks and sort/sign conventions follow the Rust text.
 * Two deliberate, documented differences:
 *   - GEMMs (ndarray::dot -> matrixmultiply 0.3.11 in the reference) are
 *     plain i-k-j loops here, so the summation order differs and results
 *     agree to rounding, not bit for bit.
 *   - Random inputs come from the caller (SplitMix64/Box-Muller in the
 *     Python test helpers), not from rand 0.8's ChaCha12 StdRng.
 * Compile with -ffp-contract=off (Rust never contracts a*b+c into an FMA).
 *
 * Storage: row-major (C order) double arrays, like ndarray's default
 * Array2<f64> layout.
 */
#include <math.h>
#include <stdio.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define LSO_POLAR_MAX_ITERS 120 /* src/lie_svd_small.rs:67 */
#define LSO_POLAR_TOL 1e-13     /* src/lie_svd_small.rs:68 */
#define LSO_JACOBI_TOL 1e-13    /* src/lie_svd_small.rs:69 */
#define LSO_MICRO_TOL 1e-13     /* src/lie_svd_micro.rs:15 */

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct {
    int polar_ok;      /* newton_schulz_polar's converged flag */
    int ns_iters;      /* Newton-Schulz iterations taken */
    int jacobi_sweeps; /* eigh_jacobi_full sweeps that ran rotations */
    int repaired;      /* repair_left_basis_from_right was used */
    int micro_fallback;/* LieSvdMicro escalated to LieSvdSmall */
    int rotations;     /* rotations applied (Jacobi / micro) */
} lso_info;

/* ------------------------------------------------------------------ */
/* dense helpers                                                       */
/* ------------------------------------------------------------------ */

static double* dalloc(size_t n) { return (double*)calloc(n ? n : 1, sizeof(double)); }

/* src/lie_svd_small.rs:71-73 */
static double frob(const double* a, size_t len) {
    double s = 0.0;
    for (size_t i = 0; i < len; ++i) s += a[i] * a[i];
    return sqrt(s);
}

/* C[m x n] = op(A) * op(B); ta/tb select transposes.  A is stored as
 * (ta ? k x m : m x k), B as (tb ? n x k : k x n), all row-major. */
static void gemm(int ta, int tb, int m, int n, int k, const double* a, const double* b, double* c) {
    for (int i = 0; i < m * n; ++i) c[i] = 0.0;
    for (int i = 0; i < m; ++i) {
        for (int p = 0; p < k; ++p) {
            double aip = ta ? a[(size_t)p * m + i] : a[(size_t)i * k + p];
            if (tb) {
                for (int j = 0; j < n; ++j) c[(size_t)i * n + j] += aip * b[(size_t)j * k + p];
            } else {
                const double* brow = b + (size_t)p * n;
                double* crow = c + (size_t)i * n;
                for (int j = 0; j < n; ++j) crow[j] += aip * brow[j];
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* LieSvdSmall                                                         */
/* ------------------------------------------------------------------ */

/* gram_schmidt_orthonormal, src/lie_svd_small.rs:82-144.  a, out: n x n. */
void lso_gram_schmidt_orthonormal(const double* a, int n, double* out) {
    double* v = dalloc(n);
    double* cand = dalloc(n);
    double* best = dalloc(n);
    memset(out, 0, sizeof(double) * (size_t)n * n);
    for (int j = 0; j < n; ++j) {
        for (int i = 0; i < n; ++i) v[i] = a[(size_t)i * n + j];
        for (int k = 0; k < j; ++k) {
            double d = 0.0;
            for (int i = 0; i < n; ++i) d += out[(size_t)i * n + k] * v[i];
            for (int i = 0; i < n; ++i) v[i] -= d * out[(size_t)i * n + k];
        }
        double norm = frob(v, n);
        if (norm >= 1e-10) {
            for (int i = 0; i < n; ++i) v[i] /= norm;
        } else {
            double best_norm = -1.0;
            for (int k = 0; k < n; ++k) {
                for (int i = 0; i < n; ++i) cand[i] = 0.0;
                cand[k] = 1.0;
                for (int col = 0; col < j; ++col) {
                    double d = 0.0;
                    for (int i = 0; i < n; ++i) d += out[(size_t)i * n + col] * cand[i];
                    for (int i = 0; i < n; ++i) cand[i] -= d * out[(size_t)i * n + col];
                }
                double en = frob(cand, n);
                if (en > best_norm) {
                    best_norm = en;
                    memcpy(best, cand, sizeof(double) * n);
                }
            }
            if (best_norm > 0.0 && isfinite(best_norm)) {
                for (int i = 0; i < n; ++i) v[i] = best[i] / best_norm;
            } else {
                for (int i = 0; i < n; ++i) v[i] = 0.0;
                v[j] = 1.0;
            }
        }
        for (int i = 0; i < n; ++i) out[(size_t)i * n + j] = v[i];
    }
    free(v);
    free(cand);
    free(best);
}

/* orthogonality_err = ||Q^T Q - I||_F, src/lie_svd_small.rs:146-150.
 * q: rows x cols; returns the cols x cols residual. */
double lso_orthogonality_err(const double* q, int rows, int cols) {
    double* g = dalloc((size_t)cols * cols);
    gemm(1, 0, cols, cols, rows, q, q, g);
    for (int i = 0; i < cols; ++i) g[(size_t)i * cols + i] -= 1.0;
    double e = frob(g, (size_t)cols * cols);
    free(g);
    return e;
}

/* newton_schulz_polar, src/lie_svd_small.rs:158-183.  Square n x n in the
 * reference; rows >= cols is accepted here (the same iteration converges to
 * the orthonormal polar factor of a tall matrix; the Gram-Schmidt fallback
 * is square-only as in the reference).  Returns polar_ok. */
int lso_newton_schulz_polar(const double* a, int rows, int cols, double* q, int* iters_out) {
    const int n = cols;
    size_t len = (size_t)rows * cols;
    double scale = frob(a, len);
    if (scale < 1e-300) scale = 1e-300;
    double* x = dalloc(len);
    double* xn = dalloc(len);
    double* m3 = dalloc((size_t)n * n);
    for (size_t i = 0; i < len; ++i) x[i] = a[i] / scale;
    double ref_norm = sqrt((double)n);
    int ok = 0, done = 0, it;
    for (it = 0; it < LSO_POLAR_MAX_ITERS; ++it) {
        /* xtx = x^T x ; x_next = x (3I - xtx) * 0.5   (:166-167) */
        gemm(1, 0, n, n, rows, x, x, m3);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) m3[(size_t)i * n + j] = (i == j ? 3.0 : 0.0) - m3[(size_t)i * n + j];
        gemm(0, 0, rows, n, n, x, m3, xn);
        int finite = 1;
        double dsq = 0.0;
        for (size_t i = 0; i < len; ++i) {
            xn[i] *= 0.5;
            if (!isfinite(xn[i])) finite = 0;
            double d = xn[i] - x[i];
            dsq += d * d;
        }
        double delta = sqrt(dsq);
        if (!finite || delta > 1e6 * ref_norm) { /* :169-171 */
            if (rows == cols) lso_gram_schmidt_orthonormal(a, n, q);
            else memcpy(q, x, sizeof(double) * len);
            if (iters_out) *iters_out = it + 1;
            free(x); free(xn); free(m3);
            return 0;
        }
        double* t = x; x = xn; xn = t;
        if (delta < LSO_POLAR_TOL * ref_norm) { /* :173-176 */
            ok = lso_orthogonality_err(x, rows, cols) <= 1e-8 * ref_norm;
            done = 1;
            ++it;
            break;
        }
    }
    if (!done) ok = lso_orthogonality_err(x, rows, cols) <= 1e-8 * ref_norm; /* :178-182 */
    memcpy(q, x, sizeof(double) * len);
    if (iters_out) *iters_out = it;
    free(x); free(xn); free(m3);
    return ok;
}

/* repair_left_basis_from_right, src/lie_svd_small.rs:185-208. */
static void repair_left_basis_from_right(const double* mat, const double* v, const double* sigma, int n,
                                         double* u_out) {
    double smax = 0.0;
    for (int j = 0; j < n; ++j) smax = fmax(smax, sigma[j]);
    double tol = 1e-12 * fmax(smax, 1.0);
    double* seed = dalloc((size_t)n * n);
    for (int j = 0; j < n; ++j) {
        if (sigma[j] > tol) {
            for (int row = 0; row < n; ++row) {
                double s = 0.0;
                for (int k = 0; k < n; ++k) s += mat[(size_t)row * n + k] * v[(size_t)k * n + j];
                seed[(size_t)row * n + j] = s / sigma[j];
            }
        }
    }
    lso_gram_schmidt_orthonormal(seed, n, u_out);
    free(seed);
}

/* Stable index sort by key descending (Rust's sort_by is stable). */
static void argsort_desc(const double* key, int n, int* order) {
    for (int i = 0; i < n; ++i) order[i] = i;
    for (int i = 1; i < n; ++i) { /* insertion sort: stable */
        int cur = order[i];
        int j = i - 1;
        while (j >= 0 && key[order[j]] < key[cur]) {
            order[j + 1] = order[j];
            --j;
        }
        order[j + 1] = cur;
    }
}

/* eigh_jacobi_full, src/lie_svd_small.rs:212-285: cyclic-by-rows two-sided
 * Jacobi with the Rutishauser tangent; eigenvalues sorted descending. */
void lso_eigh_jacobi_full(const double* p, int n, double* v_out, double* eig_out, lso_info* info) {
    double* a = dalloc((size_t)n * n);
    double* v = dalloc((size_t)n * n);
    memcpy(a, p, sizeof(double) * (size_t)n * n);
    for (int i = 0; i < n; ++i) v[(size_t)i * n + i] = 1.0;
    int max_sweeps = 60 + 20 * n;
    double ref_norm = frob(p, (size_t)n * n);
    if (ref_norm < 1e-300) ref_norm = 1e-300;
    int sweeps = 0, rotations = 0;
    for (int sw = 0; sw < max_sweeps; ++sw) {
        double off_sq = 0.0;
        for (int pp = 0; pp < n; ++pp)
            for (int qq = pp + 1; qq < n; ++qq) off_sq += a[(size_t)pp * n + qq] * a[(size_t)pp * n + qq];
        if (sqrt(off_sq) < LSO_JACOBI_TOL * ref_norm) break;
        ++sweeps;
        for (int pp = 0; pp < n; ++pp) {
            for (int qq = pp + 1; qq < n; ++qq) {
                double apq = a[(size_t)pp * n + qq];
                if (fabs(apq) < 1e-300) continue;
                double theta = (a[(size_t)qq * n + qq] - a[(size_t)pp * n + pp]) / (2.0 * apq);
                double t = theta >= 0.0 ? 1.0 / (theta + sqrt(theta * theta + 1.0))
                                        : -1.0 / (-theta + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0);
                double s = t * c;
                double app = a[(size_t)pp * n + pp];
                double aqq = a[(size_t)qq * n + qq];
                a[(size_t)pp * n + pp] = app - t * apq;
                a[(size_t)qq * n + qq] = aqq + t * apq;
                a[(size_t)pp * n + qq] = 0.0;
                a[(size_t)qq * n + pp] = 0.0;
                for (int i = 0; i < n; ++i) {
                    if (i != pp && i != qq) {
                        double aip = a[(size_t)i * n + pp];
                        double aiq = a[(size_t)i * n + qq];
                        a[(size_t)i * n + pp] = c * aip - s * aiq;
                        a[(size_t)pp * n + i] = a[(size_t)i * n + pp];
                        a[(size_t)i * n + qq] = s * aip + c * aiq;
                        a[(size_t)qq * n + i] = a[(size_t)i * n + qq];
                    }
                }
                for (int i = 0; i < n; ++i) {
                    double vip = v[(size_t)i * n + pp];
                    double viq = v[(size_t)i * n + qq];
                    v[(size_t)i * n + pp] = c * vip - s * viq;
                    v[(size_t)i * n + qq] = s * vip + c * viq;
                }
                ++rotations;
            }
        }
    }
    double* diag = dalloc(n);
    int* order = (int*)malloc(sizeof(int) * (n ? n : 1));
    for (int i = 0; i < n; ++i) diag[i] = a[(size_t)i * n + i];
    argsort_desc(diag, n, order);
    for (int nj = 0; nj < n; ++nj) {
        int oj = order[nj];
        eig_out[nj] = diag[oj];
        for (int i = 0; i < n; ++i) v_out[(size_t)i * n + nj] = v[(size_t)i * n + oj];
    }
    if (info) {
        info->jacobi_sweeps = sweeps;
        info->rotations += rotations;
    }
    free(a); free(v); free(diag); free(order);
}

/* LieSvdSmall::solve, src/lie_svd_small.rs:296-323.  u, vt: n x n; s: n. */
void lso_small_solve(const double* mat, int n, double* u, double* s, double* vt, lso_info* info) {
    lso_info local;
    memset(&local, 0, sizeof local);
    size_t nn = (size_t)n * n;
    double* q = dalloc(nn);
    double* p = dalloc(nn);
    double* v = dalloc(nn);
    double* eig = dalloc(n);
    local.polar_ok = lso_newton_schulz_polar(mat, n, n, q, &local.ns_iters);
    gemm(1, 0, n, n, n, q, mat, p); /* p = q^T mat (:305) */
    for (int i = 0; i < n; ++i) /* symmetrize (:309) */
        for (int j = i + 1; j < n; ++j) {
            double m = (p[(size_t)i * n + j] + p[(size_t)j * n + i]) * 0.5;
            p[(size_t)i * n + j] = m;
            p[(size_t)j * n + i] = m;
        }
    lso_eigh_jacobi_full(p, n, v, eig, &local);
    for (int j = 0; j < n; ++j) s[j] = fmax(eig[j], 0.0); /* :313-316 (max(0)) */
    gemm(0, 0, n, n, n, q, v, u); /* u = q v (:317) */
    if (!local.polar_ok || lso_orthogonality_err(u, n, n) > 1e-8 * sqrt((double)n)) { /* :318-320 */
        repair_left_basis_from_right(mat, v, s, n, u);
        local.repaired = 1;
    }
    for (int i = 0; i < n; ++i) /* vt = v^T (:321) */
        for (int j = 0; j < n; ++j) vt[(size_t)i * n + j] = v[(size_t)j * n + i];
    if (info) *info = local;
    free(q); free(p); free(v); free(eig);
}

/* qr_reduce, src/lie_svd_small.rs:373-418: MGS with a scale-relative rank
 * floor; dependent columns keep a zero Q column and the tiny R pivot. */
void lso_qr_reduce(const double* mat, int n, int d, double* q, double* r) {
    double* v = dalloc((size_t)n * d);
    memset(q, 0, sizeof(double) * (size_t)n * d);
    memset(r, 0, sizeof(double) * (size_t)d * d);
    double mat_scale = frob(mat, (size_t)n * d);
    if (mat_scale < 1e-300) mat_scale = 1e-300;
    double floor_ = 1e-10 * mat_scale;
    for (int j = 0; j < d; ++j) {
        for (int i = 0; i < n; ++i) v[(size_t)i * d + j] = mat[(size_t)i * d + j];
        for (int k = 0; k < j; ++k) {
            double dot = 0.0;
            for (int i = 0; i < n; ++i) dot += q[(size_t)i * d + k] * v[(size_t)i * d + j];
            r[(size_t)k * d + j] = dot;
            for (int i = 0; i < n; ++i) v[(size_t)i * d + j] -= dot * q[(size_t)i * d + k];
        }
        double ns = 0.0;
        for (int i = 0; i < n; ++i) ns += v[(size_t)i * d + j] * v[(size_t)i * d + j];
        double norm = sqrt(ns);
        r[(size_t)j * d + j] = norm;
        if (norm >= floor_)
            for (int i = 0; i < n; ++i) q[(size_t)i * d + j] = v[(size_t)i * d + j] / norm;
    }
    free(v);
}

/* LieSvdSmall::solve_rectangular, src/lie_svd_small.rs:340-360.
 * mat: rows x cols.  k = min(rows, cols); u: rows x k, s: k, vt: k x cols. */
void lso_small_solve_rectangular(const double* mat, int rows, int cols, double* u, double* s, double* vt,
                                 lso_info* info) {
    if (rows == cols) {
        lso_small_solve(mat, rows, u, s, vt, info);
        return;
    }
    if (cols > rows) { /* transpose branch (:350-354) */
        double* t = dalloc((size_t)rows * cols);
        for (int i = 0; i < rows; ++i)
            for (int j = 0; j < cols; ++j) t[(size_t)j * rows + i] = mat[(size_t)i * cols + j];
        int k = rows;
        double* ut = dalloc((size_t)cols * k); /* cols x k */
        double* vtt = dalloc((size_t)k * rows); /* k x rows */
        lso_small_solve_rectangular(t, cols, rows, ut, s, vtt, info);
        for (int i = 0; i < rows; ++i) /* u = vt_t^T : rows x k */
            for (int j = 0; j < k; ++j) u[(size_t)i * k + j] = vtt[(size_t)j * rows + i];
        for (int i = 0; i < k; ++i) /* vt = u_t^T : k x cols */
            for (int j = 0; j < cols; ++j) vt[(size_t)i * cols + j] = ut[(size_t)j * k + i];
        free(t); free(ut); free(vtt);
        return;
    }
    int d = cols; /* tall (:356-359) */
    double* q = dalloc((size_t)rows * d);
    double* r = dalloc((size_t)d * d);
    double* ur = dalloc((size_t)d * d);
    lso_qr_reduce(mat, rows, d, q, r);
    lso_small_solve(r, d, ur, s, vt, info);
    gemm(0, 0, rows, d, d, q, ur, u);
    free(q); free(r); free(ur);
}

/* ------------------------------------------------------------------ */
/* LieSvdMicro                                                         */
/* ------------------------------------------------------------------ */

/* wrap_angle, src/lie_svd_micro.rs:118-126 (also wrap_svd_angle,
 * src/lie_svd_joint.rs:293-301). */
static double wrap_half_pi(double a) {
    while (a > M_PI / 2) a -= M_PI;
    while (a < -M_PI / 2) a += M_PI;
    return a;
}

/* local_pair_svd_angles, src/lie_svd_micro.rs:128-139. */
static void pair_svd_angles(double a, double b, double c, double d, double* tl, double* tr) {
    double sum_angle = atan2(-(b + c), a - d);
    double diff_angle = atan2(b - c, a + d);
    *tl = wrap_half_pi(0.5 * (sum_angle + diff_angle));
    *tr = wrap_half_pi(0.5 * (sum_angle - diff_angle));
}

static double offdiag_norm(const double* w, int n) { /* src/lie_svd_micro.rs:101-112 */
    double s = 0.0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            if (i != j) s += w[i * n + j] * w[i * n + j];
    return sqrt(s);
}

/* LieSvdMicro::solve + solve_fixed_schedule + extract_sorted_svd,
 * src/lie_svd_micro.rs:21-95,141-214.  n <= 4.  Returns 0, or -1 if n > 4.
 * sweep_cap (test hook, no reference counterpart) lowers the 24-sweep limit
 * of n = 3, 4 so that the escalation to LieSvdSmall (:90-92) can be forced;
 * 0 keeps the reference's limit. */
int lso_micro_solve_capped(const double* mat, int n, int sweep_cap, double* u, double* s, double* vt,
                           lso_info* info) {
    lso_info local;
    memset(&local, 0, sizeof local);
    if (n < 0 || n > 4) return -1;
    if (n == 0) {
        if (info) *info = local;
        return 0;
    }
    if (n == 1) { /* :32-40 */
        double d = mat[0];
        u[0] = d < 0.0 ? -1.0 : 1.0;
        s[0] = fabs(d);
        vt[0] = 1.0;
        if (info) *info = local;
        return 0;
    }
    static const int p2[1][2] = {{0, 1}};
    static const int p3[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    static const int p4[6][2] = {{0, 1}, {2, 3}, {0, 2}, {1, 3}, {0, 3}, {1, 2}};
    const int(*pairs)[2] = n == 2 ? p2 : (n == 3 ? p3 : p4);
    int npairs = n == 2 ? 1 : (n == 3 ? 3 : 6);
    int max_sweeps = n == 2 ? 1 : 24;
    if (n > 2 && sweep_cap > 0 && sweep_cap < max_sweeps) max_sweeps = sweep_cap;
    double w[16], ub[16], vb[16];
    memcpy(w, mat, sizeof(double) * n * n);
    memset(ub, 0, sizeof ub);
    memset(vb, 0, sizeof vb);
    for (int i = 0; i < n; ++i) ub[i * n + i] = vb[i * n + i] = 1.0;
    double ref_norm = frob(mat, (size_t)n * n);
    if (ref_norm < 1e-300) ref_norm = 1e-300;
    double tol = LSO_MICRO_TOL * fmax(ref_norm, 1.0);
    for (int sw = 0; sw < (max_sweeps > 1 ? max_sweeps : 1); ++sw) {
        double energy = offdiag_norm(w, n);
        if (energy <= tol || !isfinite(energy)) break;
        int changed = 0;
        for (int pi = 0; pi < npairs; ++pi) {
            int i = pairs[pi][0], j = pairs[pi][1];
            if (fabs(w[i * n + j]) + fabs(w[j * n + i]) <= tol) continue;
            double tl, tr;
            pair_svd_angles(w[i * n + i], w[i * n + j], w[j * n + i], w[j * n + j], &tl, &tr);
            if (fabs(tl) + fabs(tr) <= 1e-18) continue;
            double sl = sin(tl), cl = cos(tl);
            for (int col = 0; col < n; ++col) { /* apply_left_rotor :141-162 */
                double ai = w[i * n + col], aj = w[j * n + col];
                w[i * n + col] = cl * ai - sl * aj;
                w[j * n + col] = sl * ai + cl * aj;
            }
            for (int r = 0; r < n; ++r) {
                double ui = ub[r * n + i], uj = ub[r * n + j];
                ub[r * n + i] = cl * ui - sl * uj;
                ub[r * n + j] = sl * ui + cl * uj;
            }
            double sr = sin(tr), cr = cos(tr);
            for (int r = 0; r < n; ++r) { /* apply_right_rotor :164-185 */
                double ai = w[r * n + i], aj = w[r * n + j];
                w[r * n + i] = cr * ai - sr * aj;
                w[r * n + j] = sr * ai + cr * aj;
            }
            for (int r = 0; r < n; ++r) {
                double vi = vb[r * n + i], vj = vb[r * n + j];
                vb[r * n + i] = cr * vi - sr * vj;
                vb[r * n + j] = sr * vi + cr * vj;
            }
            changed = 1;
            ++local.rotations;
        }
        if (!changed) break;
    }
    if (offdiag_norm(w, n) > 1e-10 * fmax(ref_norm, 1.0)) { /* :90-92 */
        lso_small_solve(mat, n, u, s, vt, info);
        if (info) info->micro_fallback = 1;
        return 0;
    }
    /* extract_sorted_svd :187-214, stable sort by |diag| descending */
    double key[4];
    int order[4];
    for (int i = 0; i < n; ++i) key[i] = fabs(w[i * n + i]);
    argsort_desc(key, n, order);
    for (int dst = 0; dst < n; ++dst) {
        int src = order[dst];
        double d = w[src * n + src];
        s[dst] = fabs(d);
        double sign = d < 0.0 ? -1.0 : 1.0;
        for (int r = 0; r < n; ++r) {
            u[r * n + dst] = sign * ub[r * n + src];
            vt[dst * n + r] = vb[r * n + src];
        }
    }
    if (info) *info = local;
    return 0;
}

int lso_micro_solve(const double* mat, int n, double* u, double* s, double* vt, lso_info* info) {
    return lso_micro_solve_capped(mat, n, 0, u, s, vt, info);
}

/* ------------------------------------------------------------------ */
/* LieSvdJoint (Phase-JADE)                                            */
/* ------------------------------------------------------------------ */

typedef struct {
    int max_sweeps;
    double tol;
    double min_pair_scale;
    int line_search_steps;
} lso_joint_params;

typedef struct {
    double initial_offdiag;
    double final_offdiag;
    int sweeps;
    int rotations;
    int rejected_rotations;
} lso_joint_trace;

/* JointDiagonalizationParams::for_n, src/lie_svd_joint.rs:26-35. */
lso_joint_params lso_joint_params_for_n(int n) {
    lso_joint_params p;
    p.max_sweeps = 12 + (n > 1 ? n : 1);
    p.tol = 1e-12;
    p.min_pair_scale = 1e-14;
    p.line_search_steps = 6;
    return p;
}

/* wrap_jacobi_angle, src/lie_svd_joint.rs:506-514. */
static double wrap_quarter_pi(double a) {
    while (a > M_PI / 4) a -= M_PI / 2;
    while (a < -M_PI / 4) a += M_PI / 2;
    return a;
}

/* round_robin_layer_count / layer_pairs, src/lie_svd_joint.rs:516-539.
 * Writes up to n/2 pairs (lo, hi) into out; returns the count. */
int lso_layer_count(int n) { return n % 2 == 0 ? (n > 0 ? n - 1 : 0) : n; }

int lso_layer_pairs(int n, int layer, int* out) {
    int m = n % 2 == 0 ? n : n + 1;
    int ring = m - 1;
    int cnt = 0;
    for (int k = 0; k < m / 2; ++k) {
        int a, b;
        if (k == 0) {
            a = m - 1;
            b = layer % ring;
        } else {
            a = (layer + k) % ring;
            b = (layer + ring - k) % ring;
        }
        if (a < n && b < n) {
            out[2 * cnt] = a < b ? a : b;
            out[2 * cnt + 1] = a < b ? b : a;
            ++cnt;
        }
    }
    return cnt;
}

/* apply_left_rotor / apply_right_rotor / apply_basis_rotor,
 * src/lie_svd_joint.rs:401-437 (rows x cols matrix m). */
static void rot_left(double* m, int rows, int cols, int i, int j, double th) {
    (void)rows;
    double s = sin(th), c = cos(th);
    for (int col = 0; col < cols; ++col) {
        double ai = m[(size_t)i * cols + col], aj = m[(size_t)j * cols + col];
        m[(size_t)i * cols + col] = c * ai - s * aj;
        m[(size_t)j * cols + col] = s * ai + c * aj;
    }
}
static void rot_right(double* m, int rows, int cols, int i, int j, double th) {
    double s = sin(th), c = cos(th);
    for (int r = 0; r < rows; ++r) {
        double ai = m[(size_t)r * cols + i], aj = m[(size_t)r * cols + j];
        m[(size_t)r * cols + i] = c * ai - s * aj;
        m[(size_t)r * cols + j] = s * ai + c * aj;
    }
}

static double offdiag_sq(const double* m, int rows, int cols) { /* :461-473 */
    double s = 0.0;
    for (int i = 0; i < rows; ++i)
        for (int j = 0; j < cols; ++j)
            if (i != j) s += m[(size_t)i * cols + j] * m[(size_t)i * cols + j];
    return s;
}

static double local_offdiag_sq(const double* m, int rows, int cols, int i, int j) { /* :475-494 */
    double s = 0.0;
    for (int col = 0; col < cols; ++col) {
        if (col != i) s += m[(size_t)i * cols + col] * m[(size_t)i * cols + col];
        if (col != j) s += m[(size_t)j * cols + col] * m[(size_t)j * cols + col];
    }
    for (int r = 0; r < rows; ++r) {
        if (r != i && r != j) {
            s += m[(size_t)r * cols + i] * m[(size_t)r * cols + i];
            s += m[(size_t)r * cols + j] * m[(size_t)r * cols + j];
        }
    }
    return s;
}

static double joint_offdiag_norm(const double* w, int k, int rows, int cols) { /* :439-441 */
    double s = 0.0;
    for (int q = 0; q < k; ++q) s += offdiag_sq(w + (size_t)q * rows * cols, rows, cols);
    return sqrt(s);
}

/* joint_pair_energy_after, src/lie_svd_joint.rs:388-399. */
static double pair_energy_after(const double* w, int k, int n, int i, int j, double th) {
    double s = sin(th), c = cos(th);
    double out = 0.0;
    for (int q = 0; q < k; ++q) {
        const double* m = w + (size_t)q * n * n;
        double a = m[(size_t)i * n + i];
        double b = 0.5 * (m[(size_t)i * n + j] + m[(size_t)j * n + i]);
        double d = m[(size_t)j * n + j];
        double off = 0.5 * (a - d) * (2.0 * s * c) + b * (c * c - s * s);
        out += 2.0 * off * off;
    }
    return out;
}

/* joint_symmetric_pair_angle, src/lie_svd_joint.rs:338-357. */
static double joint_sym_angle(const double* w, int k, int n, int i, int j) {
    double sxx = 0.0, syy = 0.0, sxy = 0.0;
    for (int q = 0; q < k; ++q) {
        const double* m = w + (size_t)q * n * n;
        double x = 0.5 * (m[(size_t)i * n + i] - m[(size_t)j * n + j]);
        double y = 0.5 * (m[(size_t)i * n + j] + m[(size_t)j * n + i]);
        sxx += x * x;
        syy += y * y;
        sxy += x * y;
    }
    double theta = 0.25 * atan2(2.0 * sxy, syy - sxx);
    double alt = wrap_quarter_pi(theta + M_PI / 4);
    theta = wrap_quarter_pi(theta);
    return pair_energy_after(w, k, n, i, j, alt) < pair_energy_after(w, k, n, i, j, theta) ? alt : theta;
}

/* LieSvdJoint::diagonalize_symmetric_with_params, src/lie_svd_joint.rs:71-158
 * with accept_joint_rotor :359-386.  mats: k consecutive n x n matrices
 * (not modified).  v_out: n x n basis; diag_out: k x n. */
void lso_joint_diag_sym(const double* mats, int k, int n, lso_joint_params params, double* v_out,
                        double* diag_out, lso_joint_trace* trace) {
    lso_joint_trace tr;
    memset(&tr, 0, sizeof tr);
    size_t nn = (size_t)n * n;
    memset(v_out, 0, sizeof(double) * nn);
    for (int i = 0; i < n; ++i) v_out[(size_t)i * n + i] = 1.0;
    if (n <= 1) {
        for (int q = 0; q < k; ++q)
            for (int i = 0; i < n; ++i) diag_out[(size_t)q * n + i] = mats[(size_t)q * nn + (size_t)i * n + i];
        if (trace) *trace = tr;
        return;
    }
    double* w = dalloc(nn * k);
    memcpy(w, mats, sizeof(double) * nn * k);
    double ref_norm = frob(w, nn * k);
    if (ref_norm < 1e-300) ref_norm = 1e-300;
    double pair_tol = params.min_pair_scale * fmax(ref_norm, 1.0);
    double tol = params.tol * fmax(ref_norm, 1.0);
    tr.initial_offdiag = joint_offdiag_norm(w, k, n, n);
    int* pairs = (int*)malloc(sizeof(int) * (n + 1));
    int lss = params.line_search_steps > 1 ? params.line_search_steps : 1;
    for (int sw = 0; sw < params.max_sweeps; ++sw) {
        tr.sweeps = sw + 1;
        double before = joint_offdiag_norm(w, k, n, n);
        int layers = lso_layer_count(n);
        for (int layer = 0; layer < layers; ++layer) {
            int np = lso_layer_pairs(n, layer, pairs);
            for (int pi = 0; pi < np; ++pi) {
                int i = pairs[2 * pi], j = pairs[2 * pi + 1];
                double po = 0.0; /* joint_pair_offdiag :451-453 */
                for (int q = 0; q < k; ++q)
                    po += fabs(w[(size_t)q * nn + (size_t)i * n + j]) + fabs(w[(size_t)q * nn + (size_t)j * n + i]);
                if (po <= pair_tol) continue;
                double theta = joint_sym_angle(w, k, n, i, j);
                if (fabs(theta) <= 1e-18) continue;
                /* accept_joint_rotor */
                double bef = 0.0;
                for (int q = 0; q < k; ++q) bef += local_offdiag_sq(w + (size_t)q * nn, n, n, i, j);
                double slack = 1e-14 * fmax(bef, 1.0);
                double scale = 1.0;
                int accepted = 0;
                for (int ls = 0; ls < lss; ++ls) {
                    double ang = theta * scale;
                    for (int q = 0; q < k; ++q) {
                        rot_left(w + (size_t)q * nn, n, n, i, j, ang);
                        rot_right(w + (size_t)q * nn, n, n, i, j, ang);
                    }
                    double aft = 0.0;
                    for (int q = 0; q < k; ++q) aft += local_offdiag_sq(w + (size_t)q * nn, n, n, i, j);
                    if (aft <= bef + slack && isfinite(aft)) {
                        rot_right(v_out, n, n, i, j, ang); /* apply_basis_rotor */
                        accepted = 1;
                        break;
                    }
                    for (int q = 0; q < k; ++q) {
                        rot_left(w + (size_t)q * nn, n, n, i, j, -ang);
                        rot_right(w + (size_t)q * nn, n, n, i, j, -ang);
                    }
                    scale *= 0.5;
                }
                if (accepted) ++tr.rotations;
                else ++tr.rejected_rotations;
            }
        }
        double after = joint_offdiag_norm(w, k, n, n);
        if (after <= tol || after >= before * (1.0 - 1e-10)) break;
    }
    tr.final_offdiag = joint_offdiag_norm(w, k, n, n);
    for (int q = 0; q < k; ++q)
        for (int i = 0; i < n; ++i) diag_out[(size_t)q * n + i] = w[(size_t)q * nn + (size_t)i * n + i];
    if (trace) *trace = tr;
    free(w);
    free(pairs);
}

/* pair_block_norm, src/lie_svd_joint.rs:277-280. */
static double pair_block_norm(const double* m, int cols, int i, int j) {
    double a = m[(size_t)i * cols + i], b = m[(size_t)i * cols + j];
    double c = m[(size_t)j * cols + i], d = m[(size_t)j * cols + j];
    return sqrt(a * a + b * b + c * c + d * d);
}

/* LieSvdJoint::joint_svd_with_params, src/lie_svd_joint.rs:170-262, with
 * joint_svd_pair_angles :264-275 (weighted averaged 2x2 block, then the
 * two-sided angles of local_pair_svd_angles_from_values :282-289 and
 * wrap_svd_angle :291-299) and accept_joint_svd_rotor :303-336.
 * mats: k consecutive rows x cols matrices (not modified).  u_out: rows x
 * rows; sig_out: k x min(rows, cols) (|diagonal|, diagonal_corridor
 * :501-504); vt_out: cols x cols. */
void lso_joint_svd(const double* mats, int k, int rows, int cols, lso_joint_params params, double* u_out,
                   double* sig_out, double* vt_out, lso_joint_trace* trace) {
    lso_joint_trace tr;
    memset(&tr, 0, sizeof tr);
    const int kk = rows < cols ? rows : cols;
    const size_t rc = (size_t)rows * cols;
    double* ub = dalloc((size_t)rows * rows);
    double* vb = dalloc((size_t)cols * cols);
    for (int i = 0; i < rows; ++i) ub[(size_t)i * rows + i] = 1.0;
    for (int i = 0; i < cols; ++i) vb[(size_t)i * cols + i] = 1.0;
    double* w = dalloc(rc * (k > 0 ? k : 1));
    memcpy(w, mats, sizeof(double) * rc * k);
    if (kk > 1) {
        double ref_norm = frob(w, rc * k);
        if (ref_norm < 1e-300) ref_norm = 1e-300;
        double pair_tol = params.min_pair_scale * fmax(ref_norm, 1.0);
        double tol = params.tol * fmax(ref_norm, 1.0);
        tr.initial_offdiag = joint_offdiag_norm(w, k, rows, cols);
        int* pairs = (int*)malloc(sizeof(int) * (kk + 1));
        int lss = params.line_search_steps > 1 ? params.line_search_steps : 1;
        for (int sw = 0; sw < params.max_sweeps; ++sw) {
            tr.sweeps = sw + 1;
            double before = joint_offdiag_norm(w, k, rows, cols);
            int layers = lso_layer_count(kk);
            for (int layer = 0; layer < layers; ++layer) {
                int np = lso_layer_pairs(kk, layer, pairs);
                for (int pi = 0; pi < np; ++pi) {
                    int i = pairs[2 * pi], j = pairs[2 * pi + 1];
                    double po = 0.0; /* joint_pair_offdiag :451-453 */
                    for (int q = 0; q < k; ++q)
                        po += fabs(w[(size_t)q * rc + (size_t)i * cols + j]) +
                              fabs(w[(size_t)q * rc + (size_t)j * cols + i]);
                    if (po <= pair_tol) continue;
                    double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
                    for (int q = 0; q < k; ++q) {
                        const double* m = w + (size_t)q * rc;
                        double wt = fmax(pair_block_norm(m, cols, i, j), 1e-300);
                        a += m[(size_t)i * cols + i] / wt;
                        b += m[(size_t)i * cols + j] / wt;
                        c += m[(size_t)j * cols + i] / wt;
                        d += m[(size_t)j * cols + j] / wt;
                    }
                    double tl, trr;
                    pair_svd_angles(a, b, c, d, &tl, &trr);
                    if (fabs(tl) + fabs(trr) <= 1e-18) continue;
                    double bef = 0.0;
                    for (int q = 0; q < k; ++q) bef += local_offdiag_sq(w + (size_t)q * rc, rows, cols, i, j);
                    double slack = 1e-14 * fmax(bef, 1.0);
                    double scale = 1.0;
                    int accepted = 0;
                    for (int ls = 0; ls < lss; ++ls) {
                        double al = tl * scale, ar = trr * scale;
                        for (int q = 0; q < k; ++q) {
                            rot_left(w + (size_t)q * rc, rows, cols, i, j, al);
                            rot_right(w + (size_t)q * rc, rows, cols, i, j, ar);
                        }
                        double aft = 0.0;
                        for (int q = 0; q < k; ++q) aft += local_offdiag_sq(w + (size_t)q * rc, rows, cols, i, j);
                        if (aft <= bef + slack && isfinite(aft)) {
                            rot_right(ub, rows, rows, i, j, al); /* apply_basis_rotor */
                            rot_right(vb, cols, cols, i, j, ar);
                            accepted = 1;
                            break;
                        }
                        for (int q = 0; q < k; ++q) {
                            rot_right(w + (size_t)q * rc, rows, cols, i, j, -ar);
                            rot_left(w + (size_t)q * rc, rows, cols, i, j, -al);
                        }
                        scale *= 0.5;
                    }
                    if (accepted) ++tr.rotations;
                    else ++tr.rejected_rotations;
                }
            }
            double after = joint_offdiag_norm(w, k, rows, cols);
            if (after <= tol || after >= before * (1.0 - 1e-10)) break;
        }
        tr.final_offdiag = joint_offdiag_norm(w, k, rows, cols);
        free(pairs);
    }
    for (int q = 0; q < k; ++q)
        for (int i = 0; i < kk; ++i) sig_out[(size_t)q * kk + i] = fabs(w[(size_t)q * rc + (size_t)i * cols + i]);
    memcpy(u_out, ub, sizeof(double) * rows * rows);
    for (int i = 0; i < cols; ++i) /* Vt = V^T */
        for (int j = 0; j < cols; ++j) vt_out[(size_t)i * cols + j] = vb[(size_t)j * cols + i];
    if (trace) *trace = tr;
    free(w);
    free(ub);
    free(vb);
}

/* ------------------------------------------------------------------ */
/* metrics::compute, src/lib.rs:149-194                                */
/* ------------------------------------------------------------------ */

/* out[0]=rel_recon, out[1]=orth_u, out[2]=orth_v, out[3]=sigma_max_rel,
 * out[4]=sigma_tail_rel (NaN when sigma_ref is NULL).  Square n x n. */
void lso_metrics(const double* a, int n, const double* u, const double* s, const double* vt,
                 const double* sigma_ref, double* out) {
    size_t nn = (size_t)n * n;
    double* us = dalloc(nn);
    double* rec = dalloc(nn);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) us[(size_t)i * n + j] = u[(size_t)i * n + j] * s[j];
    gemm(0, 0, n, n, n, us, vt, rec);
    double num = 0.0;
    for (size_t i = 0; i < nn; ++i) num += (rec[i] - a[i]) * (rec[i] - a[i]);
    double an = frob(a, nn);
    out[0] = sqrt(num) / fmax(an, 1e-300);
    out[1] = lso_orthogonality_err(u, n, n);
    gemm(0, 1, n, n, n, vt, vt, rec); /* vt vt^T (:162) */
    for (int i = 0; i < n; ++i) rec[(size_t)i * n + i] -= 1.0;
    out[2] = frob(rec, nn);
    out[3] = out[4] = NAN;
    if (sigma_ref) {
        double* got = dalloc(n);
        double* want = dalloc(n);
        int* og = (int*)malloc(sizeof(int) * (n ? n : 1));
        int* ow = (int*)malloc(sizeof(int) * (n ? n : 1));
        argsort_desc(s, n, og);
        argsort_desc(sigma_ref, n, ow);
        for (int i = 0; i < n; ++i) {
            got[i] = s[og[i]];
            want[i] = sigma_ref[ow[i]];
        }
        double mx = 0.0, tail = 0.0;
        int tail_start = (n * 3) / 4;
        for (int i = 0; i < n; ++i) {
            double rel = fabs(got[i] - want[i]) / fmax(fabs(want[i]), 1e-300);
            mx = fmax(mx, rel);
            if (i >= tail_start) tail = fmax(tail, rel);
        }
        out[3] = mx;
        out[4] = tail;
        free(got); free(want); free(og); free(ow);
    }
    free(us);
    free(rec);
}

/* ------------------------------------------------------------------ */
/* Batch drivers for the CPU baseline (one matrix per OpenMP thread)   */
/* ------------------------------------------------------------------ */

/* Ragged batch of row-major matrices: matrix b has shape m[b] x n[b] at
 * a + a_off[b]; outputs at u + u_off[b] (m x k), s + s_off[b] (k) and
 * vt + vt_off[b] (k x n), k = min(m, n).  Square matrices with n <= 4 go
 * to LieSvdMicro, as LieSvd's router does (src/lie_svd_adaptive.rs:93);
 * everything else goes through solve_rectangular.  threads <= 0 means the
 * OpenMP default. */
void lso_svd_batch(int count, const int* m, const int* n, const int64_t* a_off, const double* a,
                   const int64_t* u_off, double* u, const int64_t* s_off, double* s, const int64_t* vt_off,
                   double* vt, lso_info* info, int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
#else
    (void)threads;
#endif
    for (int b = 0; b < count; ++b) {
        lso_info* ib = info ? info + b : NULL;
        if (m[b] == n[b] && n[b] <= 4)
            lso_micro_solve(a + a_off[b], n[b], u + u_off[b], s + s_off[b], vt + vt_off[b], ib);
        else
            lso_small_solve_rectangular(a + a_off[b], m[b], n[b], u + u_off[b], s + s_off[b], vt + vt_off[b],
                                        ib);
    }
}

/* Joint diagonalization of several independent families (one per thread). */
void lso_joint_batch(int families, int k, int n, const double* mats, lso_joint_params params, double* v,
                     double* diag, lso_joint_trace* traces, int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
#else
    (void)threads;
#endif
    for (int f = 0; f < families; ++f)
        lso_joint_diag_sym(mats + (size_t)f * k * n * n, k, n, params, v + (size_t)f * n * n,
                           diag + (size_t)f * k * n, traces + f);
}

/* NS polar iterations only (the CPU-baseline timing sample for the large
 * polar workload): runs exactly `iters` iterations of :166-167 on a rows x
 * cols matrix, in place, with an OpenMP-parallel GEMM. */
void lso_ns_iterations(double* x, int rows, int cols, int iters, int threads) {
    int n = cols;
    double* g = dalloc((size_t)n * n);
    double* xn = dalloc((size_t)rows * n);
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    for (int it = 0; it < iters; ++it) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
        for (int i = 0; i < n; ++i) {
            double* gi = g + (size_t)i * n;
            for (int j = 0; j < n; ++j) gi[j] = 0.0;
            for (int p = 0; p < rows; ++p) {
                double xpi = x[(size_t)p * n + i];
                const double* xp = x + (size_t)p * n;
                for (int j = 0; j < n; ++j) gi[j] += xpi * xp[j];
            }
            for (int j = 0; j < n; ++j) gi[j] = (i == j ? 3.0 : 0.0) - gi[j];
        }
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
        for (int i = 0; i < rows; ++i) {
            double* xi = xn + (size_t)i * n;
            for (int j = 0; j < n; ++j) xi[j] = 0.0;
            for (int p = 0; p < n; ++p) {
                double a = x[(size_t)i * n + p];
                const double* gp = g + (size_t)p * n;
                for (int j = 0; j < n; ++j) xi[j] += a * gp[j];
            }
            for (int j = 0; j < n; ++j) xi[j] *= 0.5;
        }
        memcpy(x, xn, sizeof(double) * (size_t)rows * n);
    }
    free(g);
    free(xn);
}

/* ------------------------------------------------------------------ */
/* FP64 analytic gradient of L = sum_i w_i s_i + <G, U Vt> (test oracle)  */
/* ------------------------------------------------------------------ */

/* No reference counterpart (the reference has no derivatives).  For the
 * thin SVD A = U diag(s) Vt (m >= n) from LieSvdSmall::solve_rectangular:
 *   dL/dA = U (diag(w) + K) Vt + (I - U U^T) G V diag(1/s) Vt,
 *   X = U^T G V,  K_ij = (X_ij - X_ji) / (s_i + s_j).
 * Returns the loss; grad is m x n. */
double lso_svd_loss_grad(const double* a, int m, int n, const double* w, const double* g, double* grad) {
    double* u = dalloc((size_t)m * n);
    double* s = dalloc(n);
    double* vt = dalloc((size_t)n * n);
    lso_small_solve_rectangular(a, m, n, u, s, vt, NULL);
    double* q = dalloc((size_t)m * n); /* U Vt */
    gemm(0, 0, m, n, n, u, vt, q);
    double loss = 0.0;
    for (int i = 0; i < n; ++i) loss += w[i] * s[i];
    for (size_t e = 0; e < (size_t)m * n; ++e) loss += g[e] * q[e];
    double* gv = dalloc((size_t)m * n); /* G V = G Vt^T */
    gemm(0, 1, m, n, n, g, vt, gv);
    double* x = dalloc((size_t)n * n); /* U^T G V */
    gemm(1, 0, n, n, m, u, gv, x);
    double* mid = dalloc((size_t)n * n); /* diag(w) + K */
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            mid[(size_t)i * n + j] = i == j ? w[i] : (x[(size_t)i * n + j] - x[(size_t)j * n + i]) / (s[i] + s[j]);
    /* T = U mid + (G V - U X) diag(1/s) */
    double* t = dalloc((size_t)m * n);
    double* ux = dalloc((size_t)m * n);
    gemm(0, 0, m, n, n, u, mid, t);
    gemm(0, 0, m, n, n, u, x, ux);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) t[(size_t)i * n + j] += (gv[(size_t)i * n + j] - ux[(size_t)i * n + j]) / s[j];
    gemm(0, 0, m, n, n, t, vt, grad);
    free(u); free(s); free(vt); free(q); free(gv); free(x); free(mid); free(t); free(ux);
    return loss;
}

/* ---- HardwareSchedule (src/lie_svd_compiler.rs) -------------------------- */

/* HardwareSchedule::from_orthogonal_matrix (src/lie_svd_compiler.rs:144-197).
 * Givens elimination of the strict lower triangle of a d x d row-major
 * matrix: columns j = 0..d-1 in order, rows i = j+1..d-1 in order, each
 * rotation acting on rows j and i over all d columns; an entry with
 * |r_ij| <= 1e-300 is skipped (:158-160).  Event e records (i, j, c, s,
 * theta = atan2(s, c)); its layer is e (:175-187).  signs[k] = signum(r_kk)
 * after the sweep (:189), Rust's f64::signum: +-1 from the sign bit, NaN for
 * NaN.  Returns the event count. */
int lso_givens_schedule(const double* v, int d, int* ev_i, int* ev_j, double* ev_c, double* ev_s,
                        double* ev_theta, double* signs) {
    double* r = dalloc((size_t)d * d);
    memcpy(r, v, sizeof(double) * (size_t)d * d);
    int ne = 0;
    for (int j = 0; j < d; ++j) {
        for (int i = j + 1; i < d; ++i) {
            double a = r[(size_t)j * d + j];
            double b = r[(size_t)i * d + j];
            if (fabs(b) <= 1e-300) continue;
            double norm = sqrt(a * a + b * b);
            double c = a / norm;
            double s = b / norm;
            for (int k = 0; k < d; ++k) {
                double rj = r[(size_t)j * d + k];
                double ri = r[(size_t)i * d + k];
                r[(size_t)j * d + k] = c * rj + s * ri;
                r[(size_t)i * d + k] = -s * rj + c * ri;
            }
            ev_i[ne] = i;
            ev_j[ne] = j;
            ev_c[ne] = c;
            ev_s[ne] = s;
            ev_theta[ne] = atan2(s, c);
            ++ne;
        }
    }
    for (int k = 0; k < d; ++k) {
        double x = r[(size_t)k * d + k];
        signs[k] = isnan(x) ? x : copysign(1.0, x);
    }
    free(r);
    return ne;
}

/* One event of HardwarePhaseEvent (src/lie_svd_compiler.rs:27-41). */
typedef struct {
    int layer, i, j;
    double phi_l, phi_r, theta, theta_l, theta_r;
    int has_energy_before, has_energy_after;
    double energy_before, energy_after;
    const char* source;
    const char* kind;
} lso_phase_event;

typedef struct {
    char* buf;
    size_t cap, len;
} lso_sbuf;

static void sb_put(lso_sbuf* b, const char* s, size_t n) {
    if (b->len + n < b->cap) memcpy(b->buf + b->len, s, n);
    else if (b->len < b->cap) memcpy(b->buf + b->len, s, b->cap - 1 - b->len);
    b->len += n;
}
static void sb_str(lso_sbuf* b, const char* s) { sb_put(b, s, strlen(s)); }

/* Rust's `{:.17e}` (core::fmt::LowerExp with precision 17): the exact
 * decimal rounded to 18 significant digits (as C's "%.17e"), then 'e' and
 * the exponent with a '-' only when negative and no zero padding; the
 * sign of -0.0 is kept; NaN -> "NaN", +-inf -> "inf" / "-inf". */
static void sb_exp17(lso_sbuf* b, double x) {
    char t[64];
    if (isnan(x)) { sb_str(b, "NaN"); return; }
    if (isinf(x)) { sb_str(b, x < 0 ? "-inf" : "inf"); return; }
    snprintf(t, sizeof t, "%.17e", x);
    char* e = strchr(t, 'e');
    int ex = atoi(e + 1);
    *e = 0;
    sb_str(b, t);
    snprintf(t, sizeof t, "e%d", ex);
    sb_str(b, t);
}

/* Rust's Display of the signs (`{sign}`, :255): +-1 print as "1" / "-1",
 * NaN as "NaN"; other values do not occur in a sign vector. */
static void sb_sign(lso_sbuf* b, double x) {
    char t[64];
    if (isnan(x)) { sb_str(b, "NaN"); return; }
    snprintf(t, sizeof t, "%.17g", x);
    sb_str(b, t);
}

/* HardwareSchedule::to_json_string (src/lie_svd_compiler.rs:213-268).
 * Writes at most cap - 1 bytes plus a NUL into buf; returns the full
 * length. */
size_t lso_schedule_json(const char* target, int channels, int n, const lso_phase_event* ev, int n_signs,
                         const double* signs, char* buf, size_t cap) {
    lso_sbuf b = {buf, cap, 0};
    char t[128];
    sb_str(&b, "{\n  \"format_version\": \"phase-schedule-v1\",\n  \"target\": \"");
    sb_str(&b, target);
    snprintf(t, sizeof t, "\",\n  \"channels\": %d,\n  \"total_events\": %d,\n  \"events\": [\n", channels, n);
    sb_str(&b, t);
    for (int e = 0; e < n; ++e) {
        const lso_phase_event* p = ev + e;
        snprintf(t, sizeof t, "    {\"layer\": %d, \"i\": %d, \"j\": %d, \"phi_l\": ", p->layer, p->i, p->j);
        sb_str(&b, t);
        sb_exp17(&b, p->phi_l);
        sb_str(&b, ", \"phi_r\": ");
        sb_exp17(&b, p->phi_r);
        sb_str(&b, ", \"theta\": ");
        sb_exp17(&b, p->theta);
        sb_str(&b, ", \"theta_l\": ");
        sb_exp17(&b, p->theta_l);
        sb_str(&b, ", \"theta_r\": ");
        sb_exp17(&b, p->theta_r);
        sb_str(&b, ", \"source\": \"");
        sb_str(&b, p->source);
        sb_str(&b, "\", \"kind\": \"");
        sb_str(&b, p->kind);
        sb_str(&b, "\"");
        if (p->has_energy_before) {
            sb_str(&b, ", \"energy_before\": ");
            sb_exp17(&b, p->energy_before);
        }
        if (p->has_energy_after) {
            sb_str(&b, ", \"energy_after\": ");
            sb_exp17(&b, p->energy_after);
        }
        sb_str(&b, e + 1 != n ? "},\n" : "}\n");
    }
    sb_str(&b, "  ]");
    if (n_signs > 0) {
        sb_str(&b, ",\n  \"diagonal_signs\": [");
        for (int k = 0; k < n_signs; ++k) {
            if (k) sb_str(&b, ", ");
            sb_sign(&b, signs[k]);
        }
        sb_str(&b, "]");
    }
    sb_str(&b, "\n}\n");
    if (cap) buf[b.len < cap ? b.len : cap - 1] = 0;
    return b.len;
}

/* ---- Digital polish (PhaseFlow / Analog / Hybrid finishing step) -------- */

/* solve_with_digital_polish's last step (src/lie_svd_phaseflow.rs:651-657,
 * src/lie_svd_analog.rs:177-183, src/lie_svd_hybrid.rs:110-115): core =
 * U0^T A V0 with V0 = Vt0^T, (Ur, sigma, Vrt) = LieSvdSmall::solve(core),
 * U = U0 Ur, Vt = Vrt Vt0.  All n x n. */
void lso_polish(const double* a, const double* u0, const double* vt0, int n, double* u, double* s, double* vt,
                lso_info* info) {
    const size_t nn = (size_t)n * n;
    double* t = dalloc(nn);
    double* core = dalloc(nn);
    double* ur = dalloc(nn);
    double* vrt = dalloc(nn);
    gemm(1, 0, n, n, n, u0, a, t);     /* U0^T A */
    gemm(0, 1, n, n, n, t, vt0, core); /* (U0^T A) Vt0^T */
    lso_small_solve(core, n, ur, s, vrt, info);
    gemm(0, 0, n, n, n, u0, ur, u);
    gemm(0, 0, n, n, n, vrt, vt0, vt);
    free(t); free(core); free(ur); free(vrt);
}

/* ---- Adaptive triage and router (src/lie_svd_adaptive.rs) ------------- */

typedef struct {
    int n;
    int route;
    double offdiag_ratio, diagonal_dominance, row_cv, col_cv, row_col_mismatch, symmetry_ratio,
        transpose_torsion_ratio, diagonal_entropy, row_mass_entropy, col_mass_entropy, phase_mean_stress,
        phase_max_twist, phase_causal_disbalance, phase_entropy_gap, phase_torsion_energy, phase_chirality_balance,
        phase_entropy, suspicious_score;
} lso_triage;

typedef struct {
    int small_max_n, coreflow_topo_max_n, phaseflow_max_n;
    double nearly_diagonal_ratio, suspicious_threshold;
} lso_adaptive_params;

/* LieSvdAdaptiveParams::default (src/lie_svd_adaptive.rs:68-78). */
lso_adaptive_params lso_adaptive_params_default(void) {
    lso_adaptive_params p = {512, 64, 64, 1e-5, 1.15};
    return p;
}

typedef struct {
    double scalar_mean, scalar_sq, vector_sq, delay_bivector_sq, gradient_sq, energy_entropy, twist_ratio, phase;
} lso_vphase;

/* analyze_vector_phase_by + normalized_energy_entropy_by
 * (src/lie_svd_phasehealth.rs:146-211, 313-329): x[k * stride], k < n. */
static lso_vphase vector_phase(const double* x, int n, int stride) {
    lso_vphase h = {0, 0, 0, 0, 0, 0, 0, 0};
    if (n == 0) return h;
    const double nf = (double)n;
    double sum = 0.0;
    for (int i = 0; i < n; ++i) sum += x[(size_t)i * stride];
    const double mean = sum / nf;
    h.scalar_mean = mean;
    h.scalar_sq = nf * mean * mean;
    double vector_sq = 0.0, norm_sq = 0.0, delay_dot = 0.0, gradient_sq = 0.0, energy_sum = 0.0;
    for (int i = 0; i < n; ++i) {
        const double xi = x[(size_t)i * stride];
        const double xn = x[(size_t)((i + 1) % n) * stride];
        const double c = xi - mean;
        vector_sq += c * c;
        norm_sq += xi * xi;
        delay_dot += xi * xn;
        const double d = xn - xi;
        gradient_sq += d * d;
        energy_sum += xi * xi;
    }
    h.vector_sq = vector_sq;
    h.gradient_sq = gradient_sq;
    h.delay_bivector_sq = fmax(norm_sq * norm_sq - delay_dot * delay_dot, 0.0);
    double ent = 0.0;
    if (n > 1 && energy_sum > 1e-300) {
        for (int i = 0; i < n; ++i) {
            const double v = x[(size_t)i * stride];
            const double p = (v * v) / energy_sum;
            if (p > 0.0) ent -= p * log(p);
        }
        ent /= log(nf);
    }
    h.energy_entropy = ent;
    h.twist_ratio = sqrt(h.delay_bivector_sq) / fmax(norm_sq, 1e-300);
    const double orient = delay_dot >= 0.0 ? 1.0 : -1.0;
    h.phase = orient * atan2(sqrt(gradient_sq), fabs(mean) + sqrt(vector_sq) + 1e-300);
    return h;
}

/* coeff_var, relative_l2_mismatch, entropy01 (src/lie_svd_adaptive.rs:348-392). */
static double coeff_var(const double* v, int n) {
    double mean = 0.0;
    for (int i = 0; i < n; ++i) mean += v[i];
    mean /= (double)(n > 1 ? n : 1);
    if (fabs(mean) <= 1e-300) return 0.0;
    double var = 0.0;
    for (int i = 0; i < n; ++i) var += (v[i] - mean) * (v[i] - mean);
    var /= (double)(n > 1 ? n : 1);
    return sqrt(var) / fabs(mean);
}
static double rel_l2_mismatch(const double* a, const double* b, int n) {
    double num = 0.0, den = 0.0;
    for (int i = 0; i < n; ++i) {
        const double d = a[i] - b[i];
        num += d * d;
        den += a[i] * a[i] + b[i] * b[i];
    }
    return sqrt(num) / fmax(sqrt(den), 1e-300);
}
static double entropy01(const double* v, int n) {
    double sum = 0.0;
    for (int i = 0; i < n; ++i) sum += v[i];
    if (sum <= 1e-300 || n <= 1) return 1.0;
    double h = 0.0;
    for (int i = 0; i < n; ++i) {
        const double p = v[i] / sum;
        if (p > 0.0) h -= p * log(p);
    }
    return h / fmax(log((double)n), 1e-300);
}

/* analyze_block4_signature's dual_balance (src/lie_svd_block4.rs:210-266). */
static double block4_dual_balance(const double* a, int rows, int cols) {
    const int n = rows < cols ? rows : cols;
    const int blocks = n / 4;
    if (blocks == 0) return 0.0;
    double sd2 = 0.0, asd2 = 0.0;
    const double r2 = 0.70710678118654752440; /* FRAC_1_SQRT_2 */
#define SKW(p, q) (0.5 * (a[(size_t)(st + (q)) * cols + st + (p)] - a[(size_t)(st + (p)) * cols + st + (q)]))
    for (int b = 0; b < blocks; ++b) {
        const int st = 4 * b;
        const double w12 = SKW(0, 1), w13 = SKW(0, 2), w14 = SKW(0, 3), w23 = SKW(1, 2), w24 = SKW(1, 3),
                     w34 = SKW(2, 3);
        const double sd[3] = {(w12 + w34) * r2, (w13 - w24) * r2, (w14 + w23) * r2};
        const double asd[3] = {(w12 - w34) * r2, (w13 + w24) * r2, (w14 - w23) * r2};
        sd2 += sd[0] * sd[0] + sd[1] * sd[1] + sd[2] * sd[2];
        asd2 += asd[0] * asd[0] + asd[1] * asd[1] + asd[2] * asd[2];
    }
#undef SKW
    const double s = sqrt(sd2), t = sqrt(asd2);
    return fabs(s - t) / fmax(s + t, 1e-300);
}

/* analyze_matrix (src/lie_svd_adaptive.rs:186-274) with phase_signature,
 * analyze_fractal_phase_health, summarize, triangular_disbalance and
 * global_phase_invariants (src/lie_svd_phasehealth.rs:100-140, 213-345);
 * square n x n. */
void lso_analyze_matrix(const double* a, int n, lso_triage* t) {
    double all_sq = 0.0, diag_sq = 0.0, off_sq = 0.0, skew_sq = 0.0, sym_sq = 0.0;
    double* diag_abs = dalloc(n > 0 ? n : 1);
    double* row_l1 = dalloc(n > 0 ? n : 1);
    double* col_l1 = dalloc(n > 0 ? n : 1);
    double* row_l2 = dalloc(n > 0 ? n : 1);
    double* col_l2 = dalloc(n > 0 ? n : 1);
    for (int i = 0; i < n; ++i) diag_abs[i] = row_l1[i] = col_l1[i] = row_l2[i] = col_l2[i] = 0.0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            const double v = a[(size_t)i * n + j], av = fabs(v);
            all_sq += v * v;
            row_l1[i] += av;
            col_l1[j] += av;
            row_l2[i] += v * v;
            col_l2[j] += v * v;
            if (i == j) {
                diag_sq += v * v;
                diag_abs[i] = av;
            } else {
                off_sq += v * v;
            }
        }
    for (int i = 0; i < n; ++i) {
        row_l2[i] = sqrt(row_l2[i]);
        col_l2[i] = sqrt(col_l2[i]);
        for (int j = i + 1; j < n; ++j) {
            const double x = a[(size_t)i * n + j], y = a[(size_t)j * n + i];
            const double sym = 0.5 * (x + y), skew = 0.5 * (x - y);
            sym_sq += 2.0 * sym * sym;
            skew_sq += 2.0 * skew * skew;
        }
    }
    const double norm = fmax(sqrt(all_sq), 1e-300);
    t->n = n;
    t->offdiag_ratio = sqrt(off_sq) / norm;
    t->diagonal_dominance = sqrt(diag_sq) / norm;
    t->row_cv = coeff_var(row_l2, n);
    t->col_cv = coeff_var(col_l2, n);
    t->row_col_mismatch = rel_l2_mismatch(row_l1, col_l1, n);
    t->symmetry_ratio = sqrt(sym_sq) / fmax(sqrt(sym_sq + skew_sq), 1e-300);
    t->transpose_torsion_ratio = sqrt(skew_sq) / fmax(sqrt(sym_sq + skew_sq), 1e-300);
    t->diagonal_entropy = entropy01(diag_abs, n);
    t->row_mass_entropy = entropy01(row_l1, n);
    t->col_mass_entropy = entropy01(col_l1, n);
    /* phase_signature: row and column phase health, summarised */
    double d_sum = 0.0, g_sum = 0.0, max_tw = 0.0, row_tw = 0.0, col_tw = 0.0, row_en = 0.0, col_en = 0.0;
    double sin_sum = 0.0, cos_sum = 0.0, w_sum = 0.0;
    for (int i = 0; i < n; ++i) { /* rows: :106-108, :219-225 */
        const lso_vphase h = vector_phase(a + (size_t)i * n, n, 1);
        d_sum += h.delay_bivector_sq;
        g_sum += h.gradient_sq;
        max_tw = fmax(max_tw, h.twist_ratio);
        row_tw += h.twist_ratio;
        row_en += h.energy_entropy;
        const double w = sqrt(h.scalar_sq + h.vector_sq);
        sin_sum += w * sin(h.phase);
        cos_sum += w * cos(h.phase);
        w_sum += w;
    }
    for (int j = 0; j < n; ++j) { /* columns: :109-111, :226-232 */
        const lso_vphase h = vector_phase(a + j, n, n);
        d_sum += h.delay_bivector_sq;
        g_sum += h.gradient_sq;
        max_tw = fmax(max_tw, h.twist_ratio);
        col_tw += h.twist_ratio;
        col_en += h.energy_entropy;
        const double w = sqrt(h.scalar_sq + h.vector_sq);
        sin_sum += w * sin(h.phase);
        cos_sum += w * cos(h.phase);
        w_sum += w;
    }
    (void)row_tw; (void)col_tw; (void)sin_sum; (void)cos_sum; (void)w_sum;
    const double cnt = (double)(2 * n > 1 ? 2 * n : 1);
    t->phase_mean_stress = (sqrt(d_sum) + sqrt(g_sum)) / cnt;
    t->phase_max_twist = max_tw;
    {
        double up = 0.0, lo = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) {
                up += a[(size_t)i * n + j] * a[(size_t)i * n + j];
                lo += a[(size_t)j * n + i] * a[(size_t)j * n + i];
            }
        t->phase_causal_disbalance = fabs(up - lo) / fmax(up + lo, 1e-300);
    }
    t->phase_entropy_gap = n > 0 ? fabs(row_en / n - col_en / n) : 0.0;
    {
        double tor = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) {
                const double d = (a[(size_t)i * n + j] - a[(size_t)j * n + i]) * 0.5;
                tor += 2.0 * d * d;
            }
        t->phase_torsion_energy = sqrt(tor);
    }
    t->phase_chirality_balance = block4_dual_balance(a, n, n);
    {
        const size_t count = (size_t)n * n;
        double ent = 0.0;
        if (count > 1 && all_sq > 1e-300) {
            double tot = 0.0;
            for (size_t e = 0; e < count; ++e) tot += a[e] * a[e];
            for (size_t e = 0; e < count; ++e) {
                const double p = (a[e] * a[e]) / tot;
                if (p > 0.0) ent -= p * log(p);
            }
            ent /= log((double)count);
        }
        t->phase_entropy = ent;
    }
    /* suspicious_score (:319-339) */
    double sc = 0.0;
    sc += fmax(t->offdiag_ratio - 0.55, 0.0) * 0.9;
    sc += fmax(0.35 - t->diagonal_dominance, 0.0) * 0.8;
    sc += fmax(t->row_cv, t->col_cv) * 1.4;
    sc += t->row_col_mismatch * 1.6;
    sc += t->transpose_torsion_ratio * 0.45;
    sc += t->offdiag_ratio * fmax(t->symmetry_ratio - 0.82, 0.0) * 1.8;
    sc += fmax(0.78 - t->diagonal_entropy, 0.0) * 1.2;
    t->suspicious_score = sc;
    free(diag_abs); free(row_l1); free(col_l1); free(row_l2); free(col_l2);
}

static int balanced_degenerate(const lso_triage* t, const lso_adaptive_params* p) { /* :175-184 */
    return t->suspicious_score >= p->suspicious_threshold + 0.25 && t->offdiag_ratio > 0.90 &&
           t->diagonal_dominance < 0.25 && t->row_cv >= 0.20 && t->col_cv >= 0.20 && t->row_cv < 0.45 &&
           t->col_cv < 0.45 && t->row_col_mismatch < 0.45;
}

int lso_should_use_phaseflow(const lso_triage* t, const lso_adaptive_params* p) { /* :121-150 */
    if (balanced_degenerate(t, p)) return 1;
    const int causal = t->phase_causal_disbalance > 0.88 && t->offdiag_ratio > 0.20 && t->phase_mean_stress > 1.0;
    const int locking = t->phase_mean_stress > 50.0 && t->phase_max_twist > 0.92 && t->row_cv < 0.42 &&
                        t->col_cv < 0.42 && t->row_col_mismatch < 0.30 && t->offdiag_ratio > 0.85 &&
                        t->diagonal_dominance < 0.35;
    const int chiral = t->phase_chirality_balance > 0.30 && t->offdiag_ratio > 0.20 &&
                       t->phase_torsion_energy > 1.0 && t->diagonal_dominance < 0.5;
    return causal || locking || chiral;
}

int lso_should_use_coreflow_topo(const lso_triage* t, const lso_adaptive_params* p) { /* :152-173 */
    if (t->symmetry_ratio > 0.96 && t->offdiag_ratio > 0.35 && t->diagonal_dominance < 0.95) return 1;
    if (balanced_degenerate(t, p)) return 1;
    const int mass = t->row_mass_entropy < 0.92 || t->col_mass_entropy < 0.92;
    const int balanced = t->row_col_mismatch < 0.22;
    const int avoids = t->transpose_torsion_ratio < 0.62 || t->diagonal_entropy < 0.62;
    const int not_rank1 = t->row_cv < 0.42 && t->col_cv < 0.42;
    return t->suspicious_score >= p->suspicious_threshold && mass && balanced && avoids && not_rank1;
}

int lso_is_balanced_degenerate(const lso_triage* t, const lso_adaptive_params* p) { return balanced_degenerate(t, p); }

/* choose_route (src/lie_svd_adaptive.rs:103-119): 0 Micro, 1 Small,
 * 2 PhaseFlow, 3 CoreFlowTopo, 4 Hybrid (MICRO_MAX_N = 4, src/lie_svd.rs:31). */
int lso_choose_route(const lso_triage* t, const lso_adaptive_params* p) {
    if (t->n <= 4) return 0;
    if (t->offdiag_ratio <= p->nearly_diagonal_ratio) return 1;
    if (t->n <= p->phaseflow_max_n && lso_should_use_phaseflow(t, p)) return 2;
    if (t->n <= p->coreflow_topo_max_n && lso_should_use_coreflow_topo(t, p)) return 3;
    if (t->n <= p->small_max_n) return 1;
    return 4;
}

/* ===================== Tensor HO-SVD (src/lie_svd_tensor.rs) ============== */

/* mode_gram (src/lie_svd_tensor.rs:73-119): G[a][b] = sum over the other two
 * indices of x[..a..] x[..b..], in the reference's loop order. */
static void tensor_mode_gram(const double* x, int n1, int n2, int n3, int mode, double* g) {
    if (mode == 0) {
        for (int a = 0; a < n1; ++a)
            for (int b = 0; b < n1; ++b) {
                double s = 0.0;
                for (int j = 0; j < n2; ++j)
                    for (int k = 0; k < n3; ++k)
                        s += x[((size_t)a * n2 + j) * n3 + k] * x[((size_t)b * n2 + j) * n3 + k];
                g[(size_t)a * n1 + b] = s;
            }
    } else if (mode == 1) {
        for (int a = 0; a < n2; ++a)
            for (int b = 0; b < n2; ++b) {
                double s = 0.0;
                for (int i = 0; i < n1; ++i)
                    for (int k = 0; k < n3; ++k)
                        s += x[((size_t)i * n2 + a) * n3 + k] * x[((size_t)i * n2 + b) * n3 + k];
                g[(size_t)a * n2 + b] = s;
            }
    } else {
        for (int a = 0; a < n3; ++a)
            for (int b = 0; b < n3; ++b) {
                double s = 0.0;
                for (int i = 0; i < n1; ++i)
                    for (int j = 0; j < n2; ++j)
                        s += x[((size_t)i * n2 + j) * n3 + a] * x[((size_t)i * n2 + j) * n3 + b];
                g[(size_t)a * n3 + b] = s;
            }
    }
}

/* rotate_tensor3 (src/lie_svd_tensor.rs:121-145): the reference's six-deep
 * sum, out[a][b][c] = sum_ijk q1[a][i] q2[b][j] q3[c][k] x[i][j][k]; q_m is
 * r_m x n_m. */
void lso_rotate_tensor3(const double* x, int n1, int n2, int n3, const double* q1, int r1, const double* q2, int r2,
                        const double* q3, int r3, double* out) {
    for (int a = 0; a < r1; ++a)
        for (int b = 0; b < r2; ++b)
            for (int c = 0; c < r3; ++c) {
                double s = 0.0;
                for (int i = 0; i < n1; ++i)
                    for (int j = 0; j < n2; ++j)
                        for (int k = 0; k < n3; ++k)
                            s += q1[(size_t)a * n1 + i] * q2[(size_t)b * n2 + j] * q3[(size_t)c * n3 + k] *
                                 x[((size_t)i * n2 + j) * n3 + k];
                out[((size_t)a * r2 + b) * r3 + c] = s;
            }
}

/* tensor_offdiag_norm (:147-160) and superdiag_mass_ratio (:162-183). */
static double tensor_offdiag_norm(const double* x, int n1, int n2, int n3) {
    double s = 0.0;
    for (int i = 0; i < n1; ++i)
        for (int j = 0; j < n2; ++j)
            for (int k = 0; k < n3; ++k)
                if (!(i == j && j == k)) {
                    const double v = x[((size_t)i * n2 + j) * n3 + k];
                    s += v * v;
                }
    return sqrt(s);
}

static double tensor_superdiag_ratio(const double* x, int n1, int n2, int n3) {
    const int n = n1 < n2 ? (n1 < n3 ? n1 : n3) : (n2 < n3 ? n2 : n3);
    double diag = 0.0, all = 0.0;
    for (size_t e = 0; e < (size_t)n1 * n2 * n3; ++e) all += x[e] * x[e];
    for (int i = 0; i < n; ++i) {
        const double v = x[((size_t)i * n2 + i) * n3 + i];
        diag += v * v;
    }
    return diag / (all > 1e-300 ? all : 1e-300);
}

/* LieSvdTensor::hosvd3 (src/lie_svd_tensor.rs:31-62).  trace[3] =
 * {initial_offdiag, final_offdiag, superdiag_mass_ratio}. */
void lso_hosvd3(const double* x, int n1, int n2, int n3, double* core, double* u1, double* u2, double* u3,
                double* trace) {
    const int dims[3] = {n1, n2, n3};
    double* us[3] = {u1, u2, u3};
    for (int m = 0; m < 3; ++m) {
        const int n = dims[m];
        double* g = (double*)malloc(sizeof(double) * n * n);
        double* s = (double*)malloc(sizeof(double) * n);
        double* vt = (double*)malloc(sizeof(double) * n * n);
        lso_info info;
        tensor_mode_gram(x, n1, n2, n3, m, g);
        lso_small_solve(g, n, us[m], s, vt, &info);
        free(g);
        free(s);
        free(vt);
    }
    double* q[3];
    for (int m = 0; m < 3; ++m) {  /* U_m^T */
        const int n = dims[m];
        q[m] = (double*)malloc(sizeof(double) * n * n);
        for (int r = 0; r < n; ++r)
            for (int c = 0; c < n; ++c) q[m][(size_t)r * n + c] = us[m][(size_t)c * n + r];
    }
    lso_rotate_tensor3(x, n1, n2, n3, q[0], n1, q[1], n2, q[2], n3, core);
    for (int m = 0; m < 3; ++m) free(q[m]);
    trace[0] = tensor_offdiag_norm(x, n1, n2, n3);
    trace[1] = tensor_offdiag_norm(core, n1, n2, n3);
    trace[2] = tensor_superdiag_ratio(core, n1, n2, n3);
}

/* ---- LieSvdBss::separate (src/lie_svd_bss.rs:58-101) ------------------- */

/* The first half, up to the joint diagonalizer: center_channels (:162-171),
 * covariance = X_c X_c^T / samples (:63), LieSvdSmall::solve (:65), the
 * floor eps max(max sigma, 1) (:66), whitening W[i][j] = u[j][i] /
 * sqrt(max(sigma_i, floor)) (:67-73), whitened = W X_c (:74) and the lagged
 * covariances (:75, :173-197): for each lag in 0 < lag < samples,
 * c[i][j] = sum_t x[i][t + lag] x[j][t] / (samples - lag), each term divided
 * as it is added, then (c + c^T) / 2; with no usable lag, x x^T / samples.
 * lagged must hold max(nlags, 1) x ch x ch; returns the family size. */
int lso_bss_whiten(const double* x, int ch, int samples, const int* lags, int nlags, double eps, double* whitening,
                   double* whitened, double* lagged, double* sigma, double* floor_out) {
    if (ch <= 0 || samples <= 1) return 0; /* the reference asserts (:61) */
    double* xc = (double*)malloc(sizeof(double) * ch * samples);
    for (int i = 0; i < ch; ++i) {
        double mean = 0.0;
        for (int t = 0; t < samples; ++t) mean += x[(size_t)i * samples + t];
        mean /= samples;
        for (int t = 0; t < samples; ++t) xc[(size_t)i * samples + t] = x[(size_t)i * samples + t] - mean;
    }
    double* cov = (double*)malloc(sizeof(double) * ch * ch);
    for (int i = 0; i < ch; ++i)
        for (int j = 0; j < ch; ++j) {
            double s = 0.0;
            for (int t = 0; t < samples; ++t) s += xc[(size_t)i * samples + t] * xc[(size_t)j * samples + t];
            cov[(size_t)i * ch + j] = s / samples;
        }
    double* u = (double*)malloc(sizeof(double) * ch * ch);
    double* vt = (double*)malloc(sizeof(double) * ch * ch);
    lso_info info;
    lso_small_solve(cov, ch, u, sigma, vt, &info);
    double smax = 0.0;
    for (int i = 0; i < ch; ++i) smax = fmax(smax, sigma[i]);
    const double floor_v = eps * fmax(smax, 1.0);
    for (int i = 0; i < ch; ++i) {
        const double inv = 1.0 / sqrt(fmax(sigma[i], floor_v));
        for (int j = 0; j < ch; ++j) whitening[(size_t)i * ch + j] = inv * u[(size_t)j * ch + i];
    }
    for (int i = 0; i < ch; ++i)
        for (int t = 0; t < samples; ++t) {
            double s = 0.0;
            for (int k = 0; k < ch; ++k) s += whitening[(size_t)i * ch + k] * xc[(size_t)k * samples + t];
            whitened[(size_t)i * samples + t] = s;
        }
    int kk = 0;
    double* c = (double*)malloc(sizeof(double) * ch * ch);
    for (int l = 0; l < nlags; ++l) {
        const int lag = lags[l];
        if (lag <= 0 || lag >= samples) continue;
        const double denom = (double)(samples - lag);
        for (int e = 0; e < ch * ch; ++e) c[e] = 0.0;
        for (int t = 0; t < samples - lag; ++t)
            for (int i = 0; i < ch; ++i)
                for (int j = 0; j < ch; ++j)
                    c[(size_t)i * ch + j] +=
                        whitened[(size_t)i * samples + t + lag] * whitened[(size_t)j * samples + t] / denom;
        double* out = lagged + (size_t)kk * ch * ch;
        for (int i = 0; i < ch; ++i)
            for (int j = 0; j < ch; ++j) out[(size_t)i * ch + j] = 0.5 * (c[(size_t)i * ch + j] + c[(size_t)j * ch + i]);
        ++kk;
    }
    if (kk == 0) {
        for (int i = 0; i < ch; ++i)
            for (int j = 0; j < ch; ++j) {
                double s = 0.0;
                for (int t = 0; t < samples; ++t)
                    s += whitened[(size_t)i * samples + t] * whitened[(size_t)j * samples + t];
                lagged[(size_t)i * ch + j] = s / samples;
            }
        kk = 1;
    }
    *floor_out = floor_v;
    free(xc);
    free(cov);
    free(u);
    free(vt);
    free(c);
    return kk;
}

/* channel_phase_coherence (:104-126) of one channel of `samples` values. */
static double bss_coherence(const double* x, int samples) {
    if (samples <= 1) return 0.0;
    double norm = 0.0, delay_dot = 0.0, grad = 0.0;
    for (int t = 0; t < samples; ++t) {
        const double a = x[t], b = x[(t + 1) % samples];
        norm += a * a;
        delay_dot += a * b;
        const double d = b - a;
        grad += d * d;
    }
    const double phase_lock = fabs(delay_dot) / fmax(norm, 1e-300);
    const double smooth = 1.0 / (1.0 + sqrt(grad / fmax(norm, 1e-300)));
    const double v = 0.5 * phase_lock + 0.5 * smooth;
    return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
}

/* The second half (:81-86): separated = V^T whitened, unmixing = V^T W,
 * and the channel coherence of the separated channels. */
void lso_bss_unmix(const double* basis, const double* whitening, const double* whitened, int ch, int samples,
                   double* unmixing, double* separated, double* coherence) {
    for (int i = 0; i < ch; ++i) {
        for (int t = 0; t < samples; ++t) {
            double s = 0.0;
            for (int k = 0; k < ch; ++k) s += basis[(size_t)k * ch + i] * whitened[(size_t)k * samples + t];
            separated[(size_t)i * samples + t] = s;
        }
        for (int j = 0; j < ch; ++j) {
            double s = 0.0;
            for (int k = 0; k < ch; ++k) s += basis[(size_t)k * ch + i] * whitening[(size_t)k * ch + j];
            unmixing[(size_t)i * ch + j] = s;
        }
        coherence[i] = bss_coherence(separated + (size_t)i * samples, samples);
    }
}
