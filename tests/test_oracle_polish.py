"""Pins the oracle's digital-polish restatement (solve_with_digital_polish's
finishing step, src/lie_svd_phaseflow.rs:651-657): with identity bases it is
LieSvdSmall::solve bit for bit; for any orthogonal U0, V0 the singular
values are A's (LAPACK), and the factors reconstruct A; with A's exact
singular bases the core is diagonal and sigma is its sorted |diagonal|."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from tests._util import assert_factors, imposed_spectrum


def haar(rng, n):
    q, r = np.linalg.qr(rng.standard_normal((n, n)))
    return q * np.sign(np.diag(r))[None, :]


@pytest.mark.parametrize("n", [1, 2, 4, 7, 16, 33])
def test_polish_identity_bases_is_small_solve(n):
    a = np.random.default_rng(40 + n).standard_normal((n, n))
    u, s, vt, _ = oracle.polish(a, np.eye(n), np.eye(n))
    u2, s2, vt2, _ = oracle.small_solve(a)
    assert np.array_equal(s, s2) and np.array_equal(u, u2) and np.array_equal(vt, vt2)


@pytest.mark.parametrize("n,seed", [(5, 1), (24, 2), (64, 3)])
def test_polish_orthogonal_bases_keep_sigma(n, seed):
    rng = np.random.default_rng(seed)
    sig = np.sort(np.exp(rng.uniform(np.log(1e-6), 0.0, n)))[::-1]
    a = imposed_spectrum(seed, n, n, sig)
    u0, v0 = haar(rng, n), haar(rng, n)
    u, s, vt, _ = oracle.polish(a, u0, v0.T)
    want = np.linalg.svd(a, compute_uv=False)
    assert np.max(np.abs(s - want)) <= 64 * n * np.finfo(float).eps * want[0]
    assert_factors(a, u, s, vt, 1e-12 * max(n / 32, 1), "polish")


def test_polish_exact_bases_diagonal_core():
    rng = np.random.default_rng(9)
    n = 12
    a = rng.standard_normal((n, n))
    uu, ss, vvt = np.linalg.svd(a)
    perm = rng.permutation(n)
    u, s, vt, info = oracle.polish(a, uu[:, perm], vvt[perm, :])
    assert np.allclose(s, ss, rtol=1e-13, atol=1e-14 * ss[0])
    assert np.max(np.abs(u @ np.diag(s) @ vt - a)) < 1e-13 * ss[0] * n
