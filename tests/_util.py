"""Shared helpers for the parity tests (inputs, tolerances, comparisons)."""
from __future__ import annotations

import numpy as np

from lie_cliffalg_analog_svd_amd import rng
from lie_cliffalg_analog_svd_amd.metrics import compute, sigma_errors, subspace_angle_clusters

EPS = np.finfo(np.float64).eps


def uniform_matrix(seed: int, m: int, n: int, lo: float = -1.0, hi: float = 1.0) -> np.ndarray:
    """U(lo, hi) entries, the reference's gen_range(lo..hi) distribution."""
    return lo + (hi - lo) * rng.uniform(seed, m * n).reshape(m, n)


def imposed_spectrum(seed: int, m: int, n: int, sigma) -> np.ndarray:
    """A = U diag(sigma) V^T with Haar U (m x k) and V (n x k)."""
    k = min(m, n)
    u = rng.haar(rng.mix_seed(seed, 0), m, k)
    v = rng.haar(rng.mix_seed(seed, 1), n, k)
    return (u * np.asarray(sigma)[None, :k]) @ v.T


def assert_sigma_close(got, want, n: int, rel_tol: float = 1e-12, abs_c: float = 64.0, rel_floor: float = 1e-3,
                       what: str = ""):
    """Two-tier sigma parity: relative error <= rel_tol for sigma_i >=
    rel_floor * sigma_max, absolute error <= abs_c * max(n, 1) * eps *
    sigma_max everywhere (the FP64 representation floor of A)."""
    e_rel, e_abs = sigma_errors(got, want, rel_floor)
    assert e_rel <= rel_tol, f"{what} sigma rel err {e_rel:.3e} > {rel_tol:.1e}"
    assert e_abs <= abs_c * max(n, 1) * EPS, f"{what} sigma abs err {e_abs:.3e} > {abs_c * max(n, 1) * EPS:.3e}"


def assert_factors(a, u, s, vt, tol: float, what: str = ""):
    """The reference's own residual checks: recon, U and V orthogonality."""
    m = compute(a, u, s, vt)
    assert np.all(np.isfinite(u)) and np.all(np.isfinite(s)) and np.all(np.isfinite(vt)), what
    assert m.rel_recon < tol, f"{what} rel_recon={m.rel_recon:.3e}"
    assert m.orth_u < tol, f"{what} orth_u={m.orth_u:.3e}"
    assert m.orth_v < tol, f"{what} orth_v={m.orth_v:.3e}"
    assert np.all(np.diff(s) <= 0), f"{what} sigma not descending"
    return m


def subspace_sine(a, b) -> float:
    """sin of the largest principal angle between span(a) and span(b)
    (orthonormal columns): ||(I - b b^T) a||_2, accurate for tiny angles."""
    r = a - b @ (b.T @ a)
    return float(np.linalg.norm(r, 2)) if r.size else 0.0


def max_subspace_angle(u_got, u_ref, sigma_ref, gap_rtol: float = 1e-6, min_rel_sigma: float = 1e-6):
    """Largest cluster subspace angle over clusters whose singular values are
    above min_rel_sigma * sigma_max (below that the subspace is set by
    rounding of A itself)."""
    s = np.asarray(sigma_ref)
    smax = float(s.max()) if s.size else 1.0
    worst = 0.0
    for i, j, ang in subspace_angle_clusters(u_got, u_ref, s, gap_rtol):
        if s[i:j].min() >= min_rel_sigma * smax:
            worst = max(worst, ang)
    return worst


def vector_parity(u_got, u_ref, s_ref, n, what):
    """Per-cluster principal angles between computed and oracle singular
    subspaces, bounded by the FP64 perturbation bound eps*||A||/gap."""
    s = np.asarray(s_ref)
    smax = max(float(s.max()), 1e-300)
    # clusters: consecutive sigma within 1e-9 sigma_max
    i = 0
    while i < s.size:
        j = i + 1
        while j < s.size and s[j - 1] - s[j] <= 1e-9 * smax:
            j += 1
        lo = s[i - 1] - s[i] if i > 0 else np.inf
        hi = s[j - 1] - s[j] if j < s.size else np.inf
        gap = min(lo, hi)
        if s[j - 1] >= 1e-8 * smax and np.isfinite(gap):
            ang = subspace_sine(u_got[:, i:j], u_ref[:, i:j])
            bound = max(1e3 * max(n, 1) * EPS * smax / gap, 1e-10)
            assert ang <= bound, f"{what}: cluster [{i},{j}) angle {ang:.3e} > {bound:.3e}"
        i = j
