"""Pins the oracle's HO-SVD (src/lie_svd_tensor.rs) on the reference's own
test and published smoke, and on what the mathematics fixes: the tensor
rotation against an independent einsum, a superdiagonal tensor (diagonal
mode Grams: U = I, core = X), the mode factors against LAPACK's eigenvectors
of the mode Grams, and invariance of |core| under orthogonal mode rotations
of the input (no GPU)."""
from __future__ import annotations

import numpy as np
import pytest

import oracle


def smoke_tensor(n, scale=5.0, lead=1.0):
    """The reference test (:189-200; 5.0 - i on the superdiagonal) and the
    stress smoke (src/bin/stress_cpu.rs:672-680; 5 / (1 + i))."""
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    off = 1e-3 * np.sin((i * 11 + j * 7 + k * 5).astype(float))
    return np.where((i == j) & (j == k), scale / (lead + i) if lead else 5.0 - i, off)


def rel(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2)) / np.sqrt(np.sum(a * a)))


def test_reference_unit_test():
    """hosvd3_reconstructs_low_rank_tensor (src/lie_svd_tensor.rs:189-200)."""
    t = smoke_tensor(5, lead=0)
    core, u1, u2, u3, tr = oracle.hosvd3(t)
    assert rel(t, oracle.rotate_tensor3(core, u1, u2, u3)) < 1e-10
    assert tr["superdiag_mass_ratio"] > 0.95


def test_published_smoke_16():
    """`stress_cpu --tensor-hosvd` at 16^3 (README.md:1414): rel. recon
    3.123e-15 (the rounding floor: 3.10e-15 here, GEMM order aside) and
    superdiagonal mass 9.999e-1."""
    t = smoke_tensor(16)
    core, u1, u2, u3, tr = oracle.hosvd3(t)
    r = rel(t, oracle.rotate_tensor3(core, u1, u2, u3))
    assert r < 1e-14 and abs(r / 3.123e-15 - 1.0) < 0.2
    assert "%.3e" % tr["superdiag_mass_ratio"] == "9.999e-01"


def test_rotate_matches_einsum():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((4, 5, 6))
    q1, q2, q3 = rng.standard_normal((3, 4)), rng.standard_normal((7, 5)), rng.standard_normal((2, 6))
    out = oracle.rotate_tensor3(x, q1, q2, q3)
    assert out.shape == (3, 7, 2)
    assert np.allclose(out, np.einsum("ai,bj,ck,ijk->abc", q1, q2, q3, x), rtol=0, atol=1e-13)


def test_superdiagonal_tensor_is_its_own_core():
    n = 7
    t = np.zeros((n, n, n))
    t[np.arange(n), np.arange(n), np.arange(n)] = 3.0 - 0.25 * np.arange(n)
    core, u1, u2, u3, tr = oracle.hosvd3(t)
    for u in (u1, u2, u3):
        assert np.max(np.abs(np.abs(u) - np.eye(n))) < 1e-14
    assert np.max(np.abs(np.abs(core) - np.abs(t))) < 1e-14
    assert tr["initial_offdiag"] == 0.0 and tr["final_offdiag"] < 1e-13 and abs(tr["superdiag_mass_ratio"] - 1) < 1e-14


@pytest.mark.parametrize("shape", [(6, 6, 6), (4, 7, 9), (9, 3, 5)])
def test_factors_are_gram_eigenvectors_and_core_invariant(shape):
    rng = np.random.default_rng(sum(shape))
    x = rng.standard_normal(shape)
    core, u1, u2, u3, tr = oracle.hosvd3(x)
    unf = [x.reshape(shape[0], -1), np.moveaxis(x, 1, 0).reshape(shape[1], -1), np.moveaxis(x, 2, 0).reshape(shape[2], -1)]
    for u, a in zip((u1, u2, u3), unf):
        w, e = np.linalg.eigh(a @ a.T)
        e = e[:, ::-1]
        assert np.max(np.abs(np.abs(u.T @ e) - np.eye(u.shape[0]))) < 1e-9
    assert rel(x, oracle.rotate_tensor3(core, u1, u2, u3)) < 1e-14
    # orthogonal mode rotations of the input leave |core| unchanged
    qs = [np.linalg.qr(rng.standard_normal((n, n)))[0] for n in shape]
    y = np.einsum("ai,bj,ck,ijk->abc", qs[0], qs[1], qs[2], x)
    core2 = oracle.hosvd3(y)[0]
    assert np.max(np.abs(np.abs(core2) - np.abs(core))) < 1e-10 * np.abs(core).max()
