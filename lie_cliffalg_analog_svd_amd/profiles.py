"""Input profiles of the reference's stress harness (src/lib.rs:197-388).

``generate(n, profile, seed)`` returns a ``MatrixCase`` with the matrix and,
where the profile imposes one, the exact singular values.  The four
deterministic profiles (JordanDefective, SparseStructured, NearlyDiagonal,
KronStructured) are reproduced bit for bit from the reference text; the
random ones use this project's SplitMix64/Box-Muller generator (same
distribution as rand's StdRng/StandardNormal, not the same bits).

Also here: the reference's closed-form and named test matrices
(src/lie_svd_benchmarks.rs:106-483) used by the known-answer tests.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from enum import Enum

import numpy as np

from . import rng


class Profile(Enum):
    UniformRandom = "uniform_random"
    DegenerateSpectrum = "degenerate_spectrum"
    ExtremeIllConditioned = "extreme_ill_conditioned"
    JordanDefective = "jordan_defective"
    SparseStructured = "sparse_structured"
    NearlyDiagonal = "nearly_diagonal"
    KronStructured = "kron_structured"


ALL_PROFILES = list(Profile)


@dataclass
class MatrixCase:
    a: np.ndarray
    sigma_ref: np.ndarray | None


def random_orthogonal(n: int, seed: int) -> np.ndarray:
    """src/lib.rs:321-340: MGS of an N(0,1) matrix (columns, in place)."""
    m = rng.normal_matrix(seed, n, n)
    for j in range(n):
        for k in range(j):
            d = float(np.dot(m[:, j], m[:, k]))
            m[:, j] -= d * m[:, k]
        m[:, j] /= max(float(np.sqrt(np.dot(m[:, j], m[:, j]))), 1e-300)
    return m


def compose_from_svd(u, sigma, v):
    """src/lib.rs:342-345: U diag(sigma) V^T."""
    return (u * np.asarray(sigma)[None, :]) @ v.T


def kron_chain_product(factors):
    out = np.ones((1, 1))
    for f in factors:
        out = np.kron(out, f)
    return out


def kron_structured_matrix(n: int) -> np.ndarray:
    """src/lib.rs:347-387."""
    if n == 0:
        return np.zeros((0, 0))
    if n == 1:
        return np.full((1, 1), 2.0)
    if n & (n - 1) == 0:
        levels = n.bit_length() - 1
        factors = []
        for k in range(levels):
            t = 0.17 * (k + 1.0)
            factors.append(np.array([[1.6 + 0.09 * k, 0.25 * math.sin(t)], [-0.18 * math.cos(t), 0.9 + 0.05 * k]]))
        return kron_chain_product(factors)
    p = 1 << (n.bit_length() - 1)  # next_power_of_two(n) / 2
    out = np.zeros((n, n))
    out[:p, :p] = kron_structured_matrix(p)
    for i in range(p, n):
        out[i, i] = 1.0 + (i - p) * 0.01
    return out


def generate(n: int, profile: Profile, seed: int) -> MatrixCase:
    """profiles::generate (src/lib.rs:230-319)."""
    if profile == Profile.UniformRandom:
        return MatrixCase(rng.normal_matrix(seed, n, n), None)
    if profile == Profile.DegenerateSpectrum:
        u = random_orthogonal(n, rng.mix_seed(seed, 0))
        v = random_orthogonal(n, rng.mix_seed(seed, 1))
        sigma = np.array([100.0 if i < n // 8 else 50.0 if i < n // 4 else 1.0 if i < n // 2 else 1e-12
                          for i in range(n)])
        return MatrixCase(compose_from_svd(u, sigma, v), sigma)
    if profile == Profile.ExtremeIllConditioned:
        u = random_orthogonal(n, rng.mix_seed(seed, 0))
        v = random_orthogonal(n, rng.mix_seed(seed, 1))
        denom = float(max(n - 1, 1))
        sigma = np.array([10.0 ** (-18.0 * i / denom) for i in range(n)])
        return MatrixCase(compose_from_svd(u, sigma, v), sigma)
    if profile == Profile.JordanDefective:
        a = np.zeros((n, n))
        for i in range(n):
            a[i, i] = 1.0 - 0.5 * (i / max(n, 1))
            if i + 1 < n:
                a[i, i + 1] = 20.0
            if i + 2 < n and i % 4 == 0:
                a[i, i + 2] = -3.0
        return MatrixCase(a, None)
    if profile == Profile.SparseStructured:
        a = np.zeros((n, n))
        for i in range(n):
            a[i, i] = 4.0 + (i % 7) * 0.25
            if i + 1 < n:
                a[i, i + 1] = -1.0
                a[i + 1, i] = 0.75
            if i + 8 < n:
                a[i, i + 8] = 0.15
                a[i + 8, i] = -0.10
        return MatrixCase(a, None)
    if profile == Profile.NearlyDiagonal:
        sigma = np.array([1.0 + (n - i) / n for i in range(n)])
        a = np.zeros((n, n))
        for i in range(n):
            a[i, i] = sigma[i]
            for j in range(n):
                if i != j:
                    a[i, j] = 1e-6 * math.sin(float(i * 131 + j * 17))
        return MatrixCase(a, sigma)
    if profile == Profile.KronStructured:
        return MatrixCase(kron_structured_matrix(n), None)
    raise ValueError(profile)


# ---- named matrices (src/lie_svd_benchmarks.rs) ---------------------------

def kahan_matrix(n: int, theta: float) -> np.ndarray:
    """src/lie_svd_benchmarks.rs:106-119."""
    c, s = math.cos(theta), math.sin(theta)
    k = np.zeros((n, n))
    sp = 1.0
    for i in range(n):
        k[i, i] = sp
        k[i, i + 1:] = -c * sp
        sp *= s
    return k


def hilbert_matrix(n: int) -> np.ndarray:
    i = np.arange(n)
    return 1.0 / (i[:, None] + i[None, :] + 1.0)


def pei_matrix(n: int, alpha: float) -> np.ndarray:
    """src/lie_svd_benchmarks.rs:148-150: alpha I + J."""
    return np.ones((n, n)) + alpha * np.eye(n)


def pei_matrix_singular_values(n: int, alpha: float) -> np.ndarray:
    """src/lie_svd_benchmarks.rs:155-161: n-1 copies of alpha, one alpha+n."""
    s = np.full(n, alpha)
    if n > 0:
        s[-1] = alpha + n
    return s


def frank_matrix(n: int) -> np.ndarray:
    """src/lie_svd_benchmarks.rs:209-219."""
    a = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if j >= i:
                a[i, j] = n - j
            elif j + 1 == i:
                a[i, j] = n - i
    return a


def forsythe_matrix(n: int, lam: float, alpha: float) -> np.ndarray:
    """src/lie_svd_benchmarks.rs:230-242."""
    a = np.zeros((n, n))
    for i in range(n):
        a[i, i] = lam
        if i + 1 < n:
            a[i, i + 1] = 1.0
    if n > 0:
        a[n - 1, 0] = alpha if n > 1 else lam
    return a


def parter_matrix(n: int) -> np.ndarray:
    """src/lie_svd_benchmarks.rs:252-254."""
    i = np.arange(n, dtype=np.float64)
    return 1.0 / (i[:, None] - i[None, :] + 0.5)


def cauchy_matrix(n: int) -> np.ndarray:
    """src/lie_svd_benchmarks.rs:264-266."""
    i = np.arange(n, dtype=np.float64)
    return 1.0 / (i[:, None] + i[None, :] + 2.0)


def vandermonde_matrix(n: int) -> np.ndarray:
    """V[i, j] = (i+1)^j (src/lie_svd_benchmarks.rs:268-290)."""
    x = np.arange(1, n + 1, dtype=np.float64)
    return x[:, None] ** np.arange(n)[None, :]


def hubbard_dimer_hamiltonian(t: float, u: float) -> np.ndarray:
    """src/lie_svd_benchmarks.rs:459-470."""
    return np.array([[u, -t, -t, 0.0], [-t, 0.0, 0.0, -t], [-t, 0.0, 0.0, -t], [0.0, -t, -t, u]])


def hubbard_dimer_eigenvalues(t: float, u: float) -> np.ndarray:
    """src/lie_svd_benchmarks.rs:480-483."""
    gap = math.sqrt((u / 2.0) ** 2 + 4.0 * t * t)
    return np.array([0.0, u, u / 2.0 - gap, u / 2.0 + gap])
