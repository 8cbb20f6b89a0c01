"""Pins the CPU oracle (oracle/lsvd_oracle.c) before it is trusted.

Each oracle function is checked against something a plausible mistake
(dropped term, wrong sign or index, transposed operand) would fail:
closed-form spectra from the reference's known-answer tests, LAPACK
(numpy), exact invariants, brute force on tiny inputs, and the reference's
own assertion thresholds.  No GPU needed.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from lie_cliffalg_analog_svd_amd import metrics as pmetrics
from lie_cliffalg_analog_svd_amd import profiles, rng
from tests._util import EPS, assert_factors, assert_sigma_close, imposed_spectrum, uniform_matrix


# ---- LieSvdSmall -----------------------------------------------------------

@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 9, 13, 17, 31])
def test_small_sizes_vs_lapack(n):
    """src/lie_svd_small.rs:557 sizes; sigma vs LAPACK gesdd."""
    a = uniform_matrix(100 + n, n, n)
    u, s, vt, info = oracle.small_solve(a)
    assert_factors(a, u, s, vt, 1e-12, f"n={n}")
    assert_sigma_close(s, np.linalg.svd(a, compute_uv=False), n, what=f"n={n}")
    assert info["polar_ok"] == 1


def test_small_random_64_anchor():
    """64x64 N(0,1): residuals ~1e-14 as the reference README reports
    (README.md:1117: rel_recon 1.15e-14, orth 8.8e-14); NS takes ~20 its."""
    a = rng.normal_matrix(0x5EED0001, 64, 64)
    u, s, vt, info = oracle.small_solve(a)
    m = assert_factors(a, u, s, vt, 1e-12)
    assert m.rel_recon < 1e-13 and m.orth_u < 5e-13 and m.orth_v < 5e-13
    assert 10 <= info["ns_iters"] <= 40
    assert_sigma_close(s, np.linalg.svd(a, compute_uv=False), 64)


@pytest.mark.parametrize("n", [16, 64])
def test_pei_closed_form(n):
    """src/lie_svd_benchmarks.rs:606-627: Pei matrix alpha I + J has sigma
    {alpha (n-1 times), alpha + n}."""
    alpha = 0.01
    a = profiles.pei_matrix(n, alpha)
    want = profiles.pei_matrix_singular_values(n, alpha)
    u, s, vt, _ = oracle.small_solve(a)
    m = pmetrics.compute(a, u, s, vt, want)
    assert m.orth_u < 1e-9 and m.orth_v < 1e-9 and m.rel_recon < 1e-9
    assert m.sigma_max_rel < 1e-8 and m.sigma_tail_rel < 1e-6
    # tighter than the reference's own bar: sigma_max to 1e-13, the
    # degenerate cluster to n * eps * sigma_max absolute
    assert abs(s[0] - (alpha + n)) / (alpha + n) < 1e-13
    assert np.max(np.abs(s[1:] - alpha)) < 8 * n * EPS * (alpha + n)


@pytest.mark.parametrize("n", [32, 64])
def test_degenerate_spectrum_profile(n):
    """src/lie_svd_benchmarks.rs:647-674 (imposed sigma 100 .. 1e-12)."""
    case = profiles.generate(n, profiles.Profile.DegenerateSpectrum, 909)
    u, s, vt, _ = oracle.small_solve(case.a)
    m = pmetrics.compute(case.a, u, s, vt, case.sigma_ref)
    assert m.orth_u < 1e-9 and m.orth_v < 1e-9 and m.rel_recon < 1e-9
    assert abs(s[0] - 100.0) / 100.0 < 1e-9
    assert m.sigma_max_rel < 0.05
    assert_sigma_close(s, np.linalg.svd(case.a, compute_uv=False), n, rel_tol=1e-12)


@pytest.mark.parametrize("n", [32, 64])
def test_extreme_ill_conditioned_profile(n):
    """src/lie_svd_benchmarks.rs:696-725: absolute error < 1e-9 everywhere,
    relative < 1e-8 where sigma > 1e-6."""
    case = profiles.generate(n, profiles.Profile.ExtremeIllConditioned, 909)
    u, s, vt, info = oracle.small_solve(case.a)
    assert np.all(np.isfinite(u)) and np.all(np.isfinite(vt)) and np.all(s >= 0)
    m = pmetrics.compute(case.a, u, s, vt)
    assert m.orth_u < 1e-9 and m.orth_v < 1e-9
    want = np.sort(case.sigma_ref)[::-1]
    for g, w in zip(s, want):
        assert abs(g - w) < 1e-9
        if w > 1e-6:
            assert abs(g - w) / w < 1e-8


def test_ill_conditioned_32():
    """src/lie_svd_small.rs:575-616: kappa 1e15, residuals < 1e-6."""
    n = 32
    sig = np.array([10.0 ** (-15.0 * i / (n - 1)) for i in range(n)])
    a = imposed_spectrum(42, n, n, sig)
    u, s, vt, _ = oracle.small_solve(a)
    m = pmetrics.compute(a, u, s, vt)
    assert m.rel_recon < 1e-6 and m.orth_u < 1e-6 and m.orth_v < 1e-6


def test_jordan_defective_64():
    """src/lie_svd_small.rs:619-640."""
    a = profiles.generate(64, profiles.Profile.JordanDefective, 0).a
    u, s, vt, _ = oracle.small_solve(a)
    m = pmetrics.compute(a, u, s, vt)
    assert m.rel_recon < 1e-10 and m.orth_u < 1e-8 and m.orth_v < 1e-8
    assert_sigma_close(s, np.linalg.svd(a, compute_uv=False), 64)


@pytest.mark.parametrize("name,builder", [
    ("kahan32", lambda: profiles.kahan_matrix(32, 1.2)),
    ("frank32", lambda: profiles.frank_matrix(32)),
    ("forsythe32", lambda: profiles.forsythe_matrix(32, 0.0, 1e-6)),
    ("parter32", lambda: profiles.parter_matrix(32)),
    ("cauchy10", lambda: profiles.cauchy_matrix(10)),
    ("hilbert12", lambda: profiles.hilbert_matrix(12)),
    ("vandermonde10", lambda: profiles.vandermonde_matrix(10)),
])
def test_named_matrices(name, builder):
    """src/lie_svd_benchmarks.rs:527-970 self-consistency bounds, plus
    LAPACK sigma (two-tier)."""
    a = builder()
    u, s, vt, _ = oracle.small_solve(a)
    m = pmetrics.compute(a, u, s, vt)
    assert m.orth_u < 1e-9 and m.orth_v < 1e-9 and m.rel_recon < 1e-9, (name, m)
    assert_sigma_close(s, np.linalg.svd(a, compute_uv=False), a.shape[0], rel_tol=1e-10, what=name)


def test_parter_clusters_near_pi():
    """src/lie_svd_benchmarks.rs:831-850."""
    for n in (16, 32, 64):
        _, s, _, _ = oracle.small_solve(profiles.parter_matrix(n))
        assert np.mean(np.abs(s - math.pi) < 0.05) > 0.75


def test_sigma_permutation_invariance():
    """src/phase_normalizer.rs:419/637: sigma invariant under row and
    column permutations."""
    a = uniform_matrix(7, 24, 24)
    _, s0, _, _ = oracle.small_solve(a)
    p = np.random.default_rng(3)
    _, s1, _, _ = oracle.small_solve(a[p.permutation(24)][:, p.permutation(24)])
    assert np.max(np.abs(s0 - s1)) < 64 * 24 * EPS * s0[0]


# ---- rectangular ---------------------------------------------------------

@pytest.mark.parametrize("m,n", [(30, 3), (20, 15), (4, 11), (64, 60), (100, 95), (128, 4), (9, 128)])
def test_rectangular_vs_lapack(m, n):
    """src/lie_svd_small.rs:462-495, 992-1028 shapes."""
    a = uniform_matrix(2026 + m * n, m, n)
    u, s, vt, _ = oracle.small_solve_rectangular(a)
    k = min(m, n)
    assert u.shape == (m, k) and vt.shape == (k, n)
    assert_factors(a, u, s, vt, 1e-12, f"{m}x{n}")
    assert_sigma_close(s, np.linalg.svd(a, compute_uv=False), max(m, n), what=f"{m}x{n}")


def test_rectangular_matches_square():
    """src/lie_svd_small.rs:498-514."""
    a = uniform_matrix(42, 12, 12)
    _, s1, _, _ = oracle.small_solve(a)
    _, s2, _, _ = oracle.small_solve_rectangular(a)
    assert np.max(np.abs(s1 - s2)) < 1e-10


def test_rectangular_rank_deficient_column():
    """src/lie_svd_small.rs:517-540: duplicate column, finite, recon < 1e-8,
    and MGS gives that column a zero Q column."""
    n = 20
    i = np.arange(n)
    a = np.zeros((n, 3))
    a[:, 0] = np.sin((i * 11 + 3) * 0.23) * 2.0
    a[:, 1] = np.cos((i * 13 + 7) * 0.41)
    a[:, 2] = a[:, 0]
    u, s, vt, _ = oracle.small_solve_rectangular(a)
    assert np.all(np.isfinite(u)) and np.all(np.isfinite(s)) and np.all(np.isfinite(vt))
    m = pmetrics.compute(a, u, s, vt)
    assert m.rel_recon < 1e-8
    q, r = oracle.qr_reduce(a)
    assert np.all(q[:, 2] == 0.0) and r[2, 2] < 1e-10 * np.linalg.norm(a)


def test_qr_reduce_invariants():
    a = uniform_matrix(5, 40, 17)
    q, r = oracle.qr_reduce(a)
    assert np.allclose(q @ r, a, atol=1e-13)
    assert np.linalg.norm(q.T @ q - np.eye(17)) < 1e-13
    assert np.all(np.tril(r, -1) == 0.0) and np.all(np.diag(r) > 0)


# ---- Newton-Schulz polar, eigh ---------------------------------------------

def test_newton_schulz_polar_is_the_polar_factor():
    """Q from NS equals U V^T of LAPACK's SVD; Q^T A is symmetric PSD."""
    a = uniform_matrix(11, 33, 33)
    q, ok, iters = oracle.newton_schulz_polar(a)
    assert ok and 5 < iters < 120
    u, _, vt = np.linalg.svd(a)
    assert np.linalg.norm(q - u @ vt) < 1e-12
    h = q.T @ a
    assert np.linalg.norm(h - h.T) < 1e-12 * np.linalg.norm(a)
    assert np.linalg.eigvalsh(0.5 * (h + h.T)).min() > -1e-12


def test_newton_schulz_tall_polar():
    a = uniform_matrix(12, 50, 20)
    q, ok, _ = oracle.newton_schulz_polar(a)
    u, _, vt = np.linalg.svd(a, full_matrices=False)
    assert ok and np.linalg.norm(q - u @ vt) < 1e-12


def test_newton_schulz_divergence_falls_back_to_gram_schmidt():
    """Non-finite input: the reference returns (gram_schmidt(A), false)."""
    a = uniform_matrix(13, 6, 6)
    a[2, 3] = np.nan
    q, ok, _ = oracle.newton_schulz_polar(a)
    assert not ok
    assert np.all(np.isfinite(q)) or np.isnan(q).any()  # GS of a NaN matrix stays finite by construction
    gs = oracle.gram_schmidt_orthonormal(a)
    assert np.array_equal(q, gs, equal_nan=True)


def test_gram_schmidt_never_emits_zero_column():
    """src/lie_svd_small.rs:75-144: rank-deficient input still gives an
    orthogonal matrix (best standard-basis fallback)."""
    a = np.zeros((5, 5))
    a[:, 0] = 1.0
    a[:, 1] = 2.0  # dependent
    q = oracle.gram_schmidt_orthonormal(a)
    assert np.linalg.norm(q.T @ q - np.eye(5)) < 1e-14


@pytest.mark.parametrize("t,u", [(1.0, 4.0), (1.0, 0.0), (2.3, -3.1), (0.5, 10.0)])
def test_hubbard_dimer_closed_form(t, u):
    """src/lie_svd_benchmarks.rs:1267-1287 (tighter: 1e-13 instead of 1e-9)."""
    h = profiles.hubbard_dimer_hamiltonian(t, u)
    v, e, _ = oracle.eigh_jacobi_full(h)
    want = np.sort(profiles.hubbard_dimer_eigenvalues(t, u))[::-1]
    assert np.all(np.diff(e) <= 0)
    assert np.max(np.abs(e - want)) < 1e-13 * max(1.0, abs(u), abs(t))
    assert np.linalg.norm(v.T @ h @ v - np.diag(e)) < 1e-13 * np.linalg.norm(h)


def test_hubbard_near_degenerate_gap():
    """src/lie_svd_benchmarks.rs:1299-1329."""
    u = 1e-12
    _, e, _ = oracle.eigh_jacobi_full(profiles.hubbard_dimer_hamiltonian(0.7, u))
    near = sorted(x for x in e if abs(x) < 1e-6)
    assert len(near) == 2
    assert abs(near[0]) < 1e-9 and abs(near[1] - u) / u < 1e-2


def test_eigh_vs_lapack_and_permutation():
    """eigh_jacobi_full vs numpy eigh; eigenvalues invariant under a
    symmetric permutation (src/phase_normalizer.rs:790)."""
    g = rng.normal_matrix(21, 30, 30)
    p = g + g.T
    v, e, _ = oracle.eigh_jacobi_full(p)
    assert np.max(np.abs(e - np.linalg.eigvalsh(p)[::-1])) < 64 * 30 * EPS * np.abs(e).max()
    assert np.linalg.norm(v.T @ v - np.eye(30)) < 1e-13
    perm = np.random.default_rng(1).permutation(30)
    _, e2, _ = oracle.eigh_jacobi_full(p[perm][:, perm])
    assert np.max(np.abs(e - e2)) < 64 * 30 * EPS * np.abs(e).max()


# ---- LieSvdMicro -------------------------------------------------------------

def _svd2_closed_form(a):
    """Textbook 2x2 singular values: (|p| +- |q|) with p, q the
    conformal/anticonformal parts."""
    e, f = (a[0, 0] + a[1, 1]) / 2, (a[0, 0] - a[1, 1]) / 2
    g, h = (a[1, 0] + a[0, 1]) / 2, (a[1, 0] - a[0, 1]) / 2
    q, r = math.hypot(e, h), math.hypot(f, g)
    return np.array([q + r, abs(q - r)])


@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_micro_random(n):
    """src/lie_svd_micro.rs:239-251 (32 matrices per n, U(-2, 2))."""
    for b in range(32):
        a = uniform_matrix(900 + n * 1000 + b, n, n, -2.0, 2.0)
        u, s, vt, info = oracle.micro_solve(a)
        assert_factors(a, u, s, vt, 1e-12, f"n={n} b={b}")
        assert_sigma_close(s, np.linalg.svd(a, compute_uv=False), n, what=f"n={n}")
        if n == 2:
            assert np.max(np.abs(s - _svd2_closed_form(a))) < 8 * EPS * s[0]
            assert info["rotations"] == 1


def test_micro_degenerate_4():
    """src/lie_svd_micro.rs:254-267."""
    a = np.zeros((4, 4))
    a[0, 0] = a[1, 1] = 10.0
    a[2, 2] = a[3, 3] = 1e-12
    a[0, 1] = 0.25
    a[2, 3] = -0.5
    u, s, vt, _ = oracle.micro_solve(a)
    assert_factors(a, u, s, vt, 1e-10)


def test_micro_sign_in_u_and_diagonal_input():
    """n=1 keeps the sign in U (:32-40); a diagonal input needs no rotation
    and sorts by |d| with the sign in U (:187-214)."""
    u, s, vt, _ = oracle.micro_solve(np.array([[-3.0]]))
    assert u[0, 0] == -1.0 and s[0] == 3.0 and vt[0, 0] == 1.0
    a = np.diag([1.0, -5.0, 2.0])
    u, s, vt, info = oracle.micro_solve(a)
    assert info["rotations"] == 0
    assert np.array_equal(s, [5.0, 2.0, 1.0])
    assert np.array_equal(u, [[0, 0, 1], [-1, 0, 0], [0, 1, 0]])
    assert np.array_equal(vt, [[0, 1, 0], [0, 0, 1], [1, 0, 0]])


# ---- Phase-JADE ----------------------------------------------------------------

def _family(n, k, seed):
    q = profiles.random_orthogonal(n, seed)
    mats = np.stack([q @ np.diag(1.0 + f * 0.3 + np.arange(n) * 0.17) @ q.T for f in range(k)])
    return q, mats


@pytest.mark.parametrize("n,k,seed", [(12, 5, 130), (10, 2, 131), (14, 9, 132)])
def test_joint_family(n, k, seed):
    """src/lie_svd_joint.rs:548-585, plus recovery of Q up to permutation
    and sign (the diagonals identify the permutation)."""
    q, mats = _family(n, k, seed)
    params = oracle.JointParams(40, 1e-12, 1e-14, 6)
    v, d, tr = oracle.joint_diag_sym(mats, params)
    assert tr["final_offdiag"] < tr["initial_offdiag"] * 1e-8
    assert tr["rotations"] > 0
    assert np.linalg.norm(v.T @ v - np.eye(n)) < 1e-10
    c = np.abs(q.T @ v)
    assert np.max(np.abs(c.max(axis=0) - 1.0)) < 1e-9
    assert np.array_equal(np.sort(c.argmax(axis=0)), np.arange(n))
    for f in range(k):
        assert np.allclose(np.sort(d[f]), np.sort(1.0 + f * 0.3 + np.arange(n) * 0.17), atol=1e-10)


def test_joint_already_diagonal():
    """src/lie_svd_joint.rs:588-596: zero rotations, off <= 1e-14."""
    mats = np.stack([np.diag(1.0 + np.arange(8) + k) for k in range(4)])
    _, _, tr = oracle.joint_diag_sym(mats)
    assert tr["final_offdiag"] <= 1e-14 and tr["rotations"] == 0


def test_joint_single_matrix_equals_eigh():
    """K=1 joint diagonalization of a symmetric matrix reproduces its
    eigenvalues (the Jacobi special case)."""
    g = rng.normal_matrix(5, 10, 10)
    p = g + g.T
    v, d, tr = oracle.joint_diag_sym(p[None], oracle.JointParams(60, 1e-14, 1e-14, 6))
    assert np.max(np.abs(np.sort(d[0]) - np.linalg.eigvalsh(p))) < 1e-11


@pytest.mark.parametrize("n", [2, 3, 6, 7, 16, 33])
def test_round_robin_schedule(n):
    """layer_pairs (src/lie_svd_joint.rs:516-539): each layer is a set of
    disjoint pairs and one sweep covers every pair exactly once."""
    seen = []
    for layer in range(oracle.layer_count(n)):
        pairs = oracle.layer_pairs(n, layer)
        idx = [i for p in pairs for i in p]
        assert len(idx) == len(set(idx)) and len(pairs) == n // 2
        assert all(0 <= a < b < n for a, b in pairs)
        seen += pairs
    assert sorted(seen) == [(a, b) for a in range(n) for b in range(a + 1, n)]


# ---- metrics: closed forms (src/lib.rs:149-194) -------------------------------

def test_metrics_closed_forms_exact_factors():
    """Exact factors of a diagonal and of a permuted diagonal matrix give
    residuals of exactly zero (every product is exact in binary64)."""
    d = np.array([5.0, 3.0, 2.0, 0.5])
    o = oracle.metrics(np.diag(d), np.eye(4), d, np.eye(4))
    assert o["rel_recon"] == 0.0 and o["orth_u"] == 0.0 and o["orth_v"] == 0.0
    p = np.eye(4)[[2, 0, 3, 1]]
    o = oracle.metrics(p @ np.diag(d), p, d, np.eye(4))
    assert o["rel_recon"] == 0.0 and o["orth_u"] == 0.0 and o["orth_v"] == 0.0
    # A = P diag(d) Q^T with Vt = Q^T a permutation too
    q = np.eye(4)[[1, 3, 0, 2]]
    o = oracle.metrics(p @ np.diag(d) @ q.T, p, d, q.T)
    assert o["rel_recon"] == 0.0 and o["orth_u"] == 0.0 and o["orth_v"] == 0.0


def test_metrics_closed_forms_perturbed_factors():
    """U = I + e E_01: U^T U - I = e (E + E^T) + e^2 E^T E has Frobenius norm
    sqrt(2 e^2 + e^4); Vt = I + f E_12 the same with f; a perturbed A gives
    rel_recon = delta / ||A||_F.  With e, f, delta powers of two every
    intermediate is exact, so the results are the correctly rounded closed
    forms.  A dropped e^2 term, a missing square root or a wrong norm would
    each change the value by far more than one ulp."""
    e, f, delta = 2.0 ** -10, 2.0 ** -7, 2.0 ** -12
    n = 3
    u = np.eye(n); u[0, 1] = e
    vt = np.eye(n); vt[1, 2] = f
    s = np.array([4.0, 2.0, 1.0])
    a = (u * s[None, :]) @ vt
    a2 = a.copy(); a2[2, 0] += delta
    o = oracle.metrics(a2, u, s, vt)
    assert o["orth_u"] == math.sqrt(2 * e * e + e ** 4)
    assert o["orth_v"] == math.sqrt(2 * f * f + f ** 4)
    assert abs(o["rel_recon"] - delta / math.sqrt(float(np.sum(a2 * a2)))) <= 2 * EPS * o["rel_recon"]
    # exact factors of the same A: zero reconstruction error
    assert oracle.metrics(a, u, s, vt)["rel_recon"] == 0.0


def test_metrics_closed_forms_sigma_errors():
    """sigma_max_rel / sigma_tail_rel: both sides sorted descending, the tail
    from floor(3n/4), the 1e-300 floor under a zero reference value."""
    n = 4
    a = np.diag([4.0, 3.0, 2.0, 1.0])
    eye = np.eye(n)
    # error only in the largest value: max = 2^-10, tail (i >= 3) = 0
    o = oracle.metrics(a, eye, np.array([4.0 * (1 + 2.0 ** -10), 3.0, 2.0, 1.0]), eye, np.array([4.0, 3.0, 2.0, 1.0]))
    assert o["sigma_max_rel"] == 2.0 ** -10 and o["sigma_tail_rel"] == 0.0
    # error only in the smallest: it is in the tail
    o = oracle.metrics(a, eye, np.array([4.0, 3.0, 2.0, 1.0 + 2.0 ** -20]), eye, np.array([4.0, 3.0, 2.0, 1.0]))
    assert o["sigma_max_rel"] == 2.0 ** -20 and o["sigma_tail_rel"] == 2.0 ** -20
    # unsorted inputs are sorted before comparing: no error
    o = oracle.metrics(a, eye, np.array([1.0, 2.0, 3.0, 4.0]), eye, np.array([4.0, 2.0, 3.0, 1.0]))
    assert o["sigma_max_rel"] == 0.0 and o["sigma_tail_rel"] == 0.0
    # a zero reference value divides by the 1e-300 floor
    o = oracle.metrics(np.diag([1.0, 0.0]), np.eye(2), np.array([1.0, 1e-310]), np.eye(2), np.array([1.0, 0.0]))
    assert o["sigma_max_rel"] == 1e-310 / 1e-300 and o["sigma_tail_rel"] == 1e-310 / 1e-300
    # n = 8: the tail starts at index 6, so an error at index 5 is not in it
    d = np.arange(8.0, 0.0, -1.0)
    got = d.copy(); got[5] *= 1 + 2.0 ** -8
    o = oracle.metrics(np.diag(d), np.eye(8), got, np.eye(8), d)
    assert o["sigma_max_rel"] == 2.0 ** -8 and o["sigma_tail_rel"] == 0.0


# ---- metrics: two independent implementations agree --------------------------

def test_metrics_match_package_restatement():
    a = uniform_matrix(3, 16, 16)
    u, s, vt, _ = oracle.small_solve(a)
    u = u + 1e-4 * uniform_matrix(4, 16, 16)  # residuals well above rounding
    vt = vt + 1e-4 * uniform_matrix(5, 16, 16)
    ref = profiles.pei_matrix_singular_values(16, 0.5)
    o = oracle.metrics(a, u, s, vt, ref)
    p = pmetrics.compute(a, u, s, vt, ref)
    for k in ("rel_recon", "orth_u", "orth_v"):
        assert abs(o[k] - getattr(p, k)) <= 1e-9 * o[k]
    assert abs(o["sigma_max_rel"] - p.sigma_max_rel) < 1e-12
    assert abs(o["sigma_tail_rel"] - p.sigma_tail_rel) < 1e-12


def test_planted_mistake_is_caught():
    """The LAPACK check above would catch a dropped symmetrization: a P
    that keeps the skew part of Q^T A gives wrong sigma when Q is only an
    approximate polar factor.  Demonstrated on the oracle's own pieces."""
    a = uniform_matrix(77, 20, 20)
    q, ok, _ = oracle.newton_schulz_polar(a)
    q_bad = q + 1e-6 * uniform_matrix(78, 20, 20)  # an inexact polar factor
    p_bad = q_bad.T @ a
    _, e_bad, _ = oracle.eigh_jacobi_full(np.triu(p_bad) + np.triu(p_bad, 1).T)  # no symmetrization
    want = np.linalg.svd(a, compute_uv=False)
    assert np.max(np.abs(e_bad - want)) > 1e-9


@pytest.mark.parametrize("m,n,seed", [(12, 8, 1), (9, 9, 2), (20, 5, 3)])
def test_svd_loss_grad_value_and_central_differences(m, n, seed):
    """The oracle's analytic gradient of L = sum w_i s_i + <G, U V^T>:
    the loss value against LAPACK (sigma and the polar factor U V^T are
    unique for distinct, nonzero singular values) and the gradient against
    central differences of the oracle's own loss along random directions."""
    rng = np.random.default_rng(seed)
    u, _ = np.linalg.qr(rng.standard_normal((m, n)))
    v, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = np.linspace(2.0, 1.0, n)
    a = (u * s) @ v.T
    w = rng.uniform(size=n)
    g = rng.standard_normal((m, n))
    loss, grad = oracle.svd_loss_grad(a, w, g)
    uu, ss, vvt = np.linalg.svd(a, full_matrices=False)
    assert abs(loss - (w @ ss + np.sum(g * (uu @ vvt)))) < 1e-12 * (1 + abs(loss))
    h = 1e-6
    for _ in range(4):
        e = rng.standard_normal((m, n))
        e /= np.linalg.norm(e)
        fd = (oracle.svd_loss_grad(a + h * e, w, g)[0] - oracle.svd_loss_grad(a - h * e, w, g)[0]) / (2 * h)
        assert abs(fd - np.sum(grad * e)) < 1e-7 * (1 + np.abs(grad).max())


def _planted_joint_svd(rows, cols, k, seed, step_f=0.2, step_i=0.11):
    """The reference tests' family: M_f = U0 Sigma_f V0^T with Haar U0, V0
    and Sigma_f[i, i] = 1 + f step_f + i step_i."""
    r = np.random.default_rng(seed)
    u0 = np.linalg.qr(r.standard_normal((rows, rows)))[0]
    v0 = np.linalg.qr(r.standard_normal((cols, cols)))[0]
    kk = min(rows, cols)
    mats = []
    for f in range(k):
        sig = np.zeros((rows, cols))
        sig[np.arange(kk), np.arange(kk)] = 1.0 + f * step_f + np.arange(kk) * step_i
        mats.append(u0 @ sig @ v0.T)
    return np.stack(mats), u0, v0


def test_joint_svd_nonsymmetric_family():
    """src/lie_svd_joint.rs:598-626: off-diagonal falls by 1e-8, U and V
    orthogonal; and the planted factors are recovered up to sign and order
    with the planted singular values."""
    mats, u0, v0 = _planted_joint_svd(10, 10, 4, 150)
    p = oracle.joint_params_for_n(10)
    p.max_sweeps = 36
    u, sg, vt, tr = oracle.joint_svd(mats, p)
    assert tr["final_offdiag"] < tr["initial_offdiag"] * 1e-8
    assert tr["rotations"] > 0 and sg.shape == (4, 10)
    assert np.linalg.norm(u.T @ u - np.eye(10)) < 1e-10
    assert np.linalg.norm(vt @ vt.T - np.eye(10)) < 1e-10
    for f in range(4):  # U^T M V is diagonal with the planted values
        d = u.T @ mats[f] @ vt.T
        assert np.linalg.norm(d - np.diag(np.diag(d))) < 1e-9
        want = np.sort(1.0 + f * 0.2 + np.arange(10) * 0.11)
        assert np.allclose(np.sort(sg[f]), want, atol=1e-9)
    c = np.abs(u0.T @ u)  # recovery up to permutation and sign
    assert np.max(np.abs(1.0 - np.max(c, axis=1))) < 1e-9


def test_joint_svd_rectangular_family():
    """src/lie_svd_joint.rs:628-660: a 9 x 13 family is accepted, the
    off-diagonal does not grow, U and V stay orthogonal."""
    mats, u0, v0 = _planted_joint_svd(9, 13, 3, 151, step_i=0.07)
    p = oracle.joint_params_for_n(9)
    p.max_sweeps = 32
    u, sg, vt, tr = oracle.joint_svd(mats, p)
    assert tr["final_offdiag"] <= tr["initial_offdiag"]
    assert sg.shape == (3, 9)
    assert np.linalg.norm(u.T @ u - np.eye(9)) < 1e-10
    assert np.linalg.norm(vt @ vt.T - np.eye(13)) < 1e-10


def test_joint_svd_single_matrix_singular_values():
    """K = 1 square: the two-sided sweeps are a Jacobi SVD, |diag| equals
    LAPACK's singular values (up to order).  Rectangular: only indices below
    min(rows, cols) are ever paired (:214-216), so the bases stay the
    identity outside the leading block."""
    r = np.random.default_rng(9)
    for n in (2, 8, 11):
        a = r.standard_normal((1, n, n))
        u, sg, vt, tr = oracle.joint_svd(a, oracle.joint_params_for_n(n))
        want = np.linalg.svd(a[0], compute_uv=False)
        assert np.allclose(np.sort(sg[0])[::-1], want, atol=1e-12 * want[0]), n
        d = u.T @ a[0] @ vt.T
        assert np.linalg.norm(d - np.diag(np.diag(d))) <= 1e-12 * np.linalg.norm(a[0]) * 10
    for rows, cols in ((7, 5), (5, 9)):
        a = r.standard_normal((2, rows, cols))
        u, sg, vt, tr = oracle.joint_svd(a)
        k = min(rows, cols)
        assert np.array_equal(u[k:, k:], np.eye(rows - k)) and not np.any(u[:k, k:]) and not np.any(u[k:, :k])
        assert np.array_equal(vt[k:, k:], np.eye(cols - k)) and not np.any(vt[:k, k:]) and not np.any(vt[k:, :k])
        assert tr["final_offdiag"] <= tr["initial_offdiag"]


def test_joint_svd_degenerate_sizes():
    """min(rows, cols) <= 1: identity bases, |diagonal|, zero trace (:183-199)."""
    a = np.array([[[-3.0, 1.0, 2.0]], [[4.0, 0.0, 1.0]]])
    u, sg, vt, tr = oracle.joint_svd(a)
    assert np.array_equal(u, np.eye(1)) and np.array_equal(vt, np.eye(3))
    assert np.array_equal(sg, np.array([[3.0], [4.0]]))
    assert tr["sweeps"] == 0 and tr["rotations"] == 0
