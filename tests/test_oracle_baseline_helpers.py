"""Pins the two oracle helpers that only the bench's CPU baselines call:
the batched Phase-JADE driver (one family per thread) against the per-family
restatement, and the fixed-count Newton-Schulz iteration against what the
mathematics fixes -- one step X <- X (3I - X^T X) / 2 keeps the singular
vectors and maps each singular value through f(s) = s (3 - s^2) / 2, and an
orthogonal matrix is a fixed point."""
from __future__ import annotations

import numpy as np

import oracle


def test_joint_batch_equals_per_family():
    rng = np.random.default_rng(21)
    fams = []
    for f in range(3):
        q, _ = np.linalg.qr(rng.standard_normal((9, 9)))
        fams.append(np.stack([(q * rng.standard_normal(9)[None, :]) @ q.T for _ in range(4)]))
    fams = np.stack(fams)
    v, d, tr = oracle.joint_batch(fams, threads=2)
    for f in range(3):
        vo, do, tro = oracle.joint_diag_sym(fams[f])
        assert np.array_equal(v[f], vo) and np.array_equal(d[f], do)
        assert tr[f] == tro


def test_ns_iterations_singular_value_map():
    rng = np.random.default_rng(22)
    m, n = 12, 7
    u, _ = np.linalg.qr(rng.standard_normal((m, n)))
    w, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = np.linspace(0.05, 1.2, n)
    x = np.ascontiguousarray((u * s[None, :]) @ w.T)
    for k in (1, 3):
        y = oracle.ns_iterations(x.copy(), k)
        want = s.copy()
        for _ in range(k):
            want = want * (3.0 - want * want) / 2.0
        got = np.linalg.svd(y, compute_uv=False)
        assert np.allclose(np.sort(got), np.sort(want), rtol=0, atol=1e-13)
        # the singular vectors stay: y = u diag(f^k(s)) w^T
        assert np.max(np.abs(y - (u * want[None, :]) @ w.T)) < 1e-13


def test_ns_iterations_orthogonal_fixed_point():
    rng = np.random.default_rng(23)
    q, _ = np.linalg.qr(rng.standard_normal((10, 10)))
    y = oracle.ns_iterations(np.ascontiguousarray(q.copy()), 5)
    assert np.max(np.abs(y - q)) < 1e-14
