"""Synthetic workloads of the benchmark configurations (inputs only).

All inputs come from the seeded SplitMix64/Box-Muller generator in
``rng``; Haar factors are QR of a Gaussian with the sign of diag(R) folded
in.  Every generator returns the matrices together with what is known
about them by construction (imposed singular values, planted bases), which
the size-independent parity checks use at full size.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import rng

SEED_SINGLE = 0x5EED0001
SEED_BATCH = 0x5EED0002
SEED_POLAR = 0x5EED0003
SEED_JADE = 0x5EED0004
SEED_VJP = 0x5EED0005


@dataclass
class Batch:
    mats: list
    sigma_ref: list            # imposed singular values (descending) or None
    kind: list = field(default_factory=list)   # "ragged" or "tiny"


def single_64x64(seed: int = SEED_SINGLE) -> np.ndarray:
    """One dense 64 x 64 matrix with iid N(0, 1) entries."""
    return rng.normal_matrix(seed, 64, 64)


def ragged_matrix(seed: int, index: int, m_range=(8, 128), kmax: float = 12.0, cluster_frac: float = 0.1):
    """One matrix of the ragged batch: m ~ U{8..128}, n ~ U{4..m},
    A = U diag(s) V^T with Haar U (m x n), V (n x n), s log-spaced from 1 to
    10^-k, k ~ U[0, kmax]; with probability cluster_frac a run of 2-4 equal
    singular values."""
    s0 = rng.mix_seed(seed, index)
    u6 = rng.uniform(s0, 6)
    m = int(m_range[0] + np.floor(u6[0] * (m_range[1] - m_range[0] + 1)))
    n = int(4 + np.floor(u6[1] * (m - 4 + 1)))
    k = u6[2] * kmax
    s = np.logspace(0.0, -k, n) if n > 1 else np.ones(1)
    if u6[3] < cluster_frac and n >= 2:
        c = min(n, 2 + int(u6[4] * 3))
        start = int(u6[5] * (n - c + 1))
        s[start:start + c] = s[start]
    u = rng.haar(rng.mix_seed(s0, 1), m, n)
    v = rng.haar(rng.mix_seed(s0, 2), n, n)
    a = (u * s[None, :]) @ v.T
    return a, np.sort(s)[::-1]


def tiny_matrix(seed: int, index: int):
    """Square 2x2 .. 4x4 with iid N(0, 1) entries."""
    s0 = rng.mix_seed(seed, 1_000_000 + index)
    n = 2 + int(rng.uniform(s0, 1)[0] * 3)
    return rng.normal_matrix(rng.mix_seed(s0, 1), n, n)


def ragged_batch(count: int = 4096, tiny: int = 8192, seed: int = SEED_BATCH, kmax: float = 12.0) -> Batch:
    """The uneven-size batch: `count` ragged tall matrices plus `tiny`
    square 2..4 matrices (for the LieSvdMicro path)."""
    mats, refs, kind = [], [], []
    for i in range(count):
        a, s = ragged_matrix(seed, i, kmax=kmax)
        mats.append(a)
        refs.append(s)
        kind.append("ragged")
    for i in range(tiny):
        mats.append(tiny_matrix(seed, i))
        refs.append(None)
        kind.append("tiny")
    return Batch(mats, refs, kind)


def pack(mats):
    """Packs row-major matrices into one f64 pool (the plan layout)."""
    return np.concatenate([np.ascontiguousarray(a, dtype=np.float64).ravel() for a in mats])
