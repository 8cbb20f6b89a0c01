"""Seeded synthetic-input generators (counter-mode SplitMix64 + Box-Muller).

The reference draws inputs from rand 0.8's StdRng (ChaCha12) with
rand_distr's ziggurat StandardNormal (src/lib.rs:230-236).  Those crates are
not available here, so inputs follow the same distributions from this
project's own generator: same distribution, not the same bits.
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int, start: int = 0) -> np.ndarray:
    """`count` outputs of SplitMix64 seeded with `seed`, from draw `start`."""
    with np.errstate(over="ignore"):
        idx = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + idx * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, count: int, start: int = 0) -> np.ndarray:
    """Uniform doubles in [0, 1) with 53 random bits."""
    return (splitmix64(seed, count, start) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, count: int) -> np.ndarray:
    """Standard normals by Box-Muller from pairs of uniforms."""
    pairs = (count + 1) // 2
    u = splitmix64(seed, 2 * pairs)
    u1 = ((u[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)  # (0, 1]
    u2 = (u[1::2] >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.empty(2 * pairs)
    z[0::2] = r * np.cos(2.0 * np.pi * u2)
    z[1::2] = r * np.sin(2.0 * np.pi * u2)
    return z[:count]


def normal_matrix(seed: int, rows: int, cols: int) -> np.ndarray:
    return normal(seed, rows * cols).reshape(rows, cols)


def haar(seed: int, n: int, cols: int | None = None) -> np.ndarray:
    """Haar-distributed orthogonal n x n (or n x cols orthonormal columns):
    QR of a Gaussian matrix with the sign of diag(R) folded into Q."""
    cols = n if cols is None else cols
    g = normal_matrix(seed, n, cols)
    q, r = np.linalg.qr(g)
    d = np.sign(np.diag(r))
    d[d == 0] = 1.0
    return q * d[None, :]


def mix_seed(base: int, index: int) -> int:
    """Distinct stream per matrix: one SplitMix64 output of (base, index)."""
    return int(splitmix64(base, 1, start=index)[0])
