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


def jade_family(seed: int, family: int, K: int = 64, n: int = 256, eps: float = 1e-6):
    """One Phase-JADE family: M_k = Q diag(d_k) Q^T + eps * sym(G_k) with a
    Haar Q, d_k ~ N(0, 1)^n and G_k iid N(0, 1).  Returns (mats (K, n, n),
    Q, d (K, n))."""
    s0 = rng.mix_seed(seed, family)
    q = rng.haar(rng.mix_seed(s0, 0), n)
    d = rng.normal(rng.mix_seed(s0, 1), K * n).reshape(K, n)
    g = rng.normal(rng.mix_seed(s0, 2), K * n * n).reshape(K, n, n)
    e = 0.5 * (g + np.transpose(g, (0, 2, 1)))
    mats = np.einsum("ij,kj,lj->kil", q, d, q, optimize=True) + eps * e
    return np.ascontiguousarray(mats), q, d


def jade_batch(families: int = 32, K: int = 64, n: int = 256, eps: float = 1e-6, seed: int = SEED_JADE):
    fams, qs, ds = [], [], []
    for f in range(families):
        m, q, d = jade_family(seed, f, K, n, eps)
        fams.append(m)
        qs.append(q)
        ds.append(d)
    return np.stack(fams), qs, ds


def joint_svd_family(n: int, K: int = 4, seed: int = SEED_JADE + 1):
    """The two-sided family of the joint-SVD stress smoke
    (src/bin/stress_cpu.rs:597-626): A_k = U0 diag(1 + 0.15 k + 0.05 i) V0^T
    with Haar U0, V0 (n x n).  Returns (mats (K, n, n), U0, V0, d (K, n))."""
    u0 = rng.haar(rng.mix_seed(seed, 2 * n), n)
    v0 = rng.haar(rng.mix_seed(seed, 2 * n + 1), n)
    i = np.arange(n, dtype=np.float64)
    d = np.stack([1.0 + 0.15 * k + 0.05 * i for k in range(K)])
    mats = np.einsum("ij,kj,lj->kil", u0, d, v0, optimize=True)
    return np.ascontiguousarray(mats), u0, v0, d


STRESS_SEED_JADE = 0x9E3779B97F4A7C15
STRESS_SEED_JOINT_SVD = 0xA24BAED4963EE407
_M64 = (1 << 64) - 1


def _seq_sum(x: np.ndarray) -> float:
    # left-to-right accumulation, one rounding per add (np.sum is pairwise)
    return float(np.cumsum(x)[-1]) if x.size else 0.0


def stress_orthogonal(n: int, state: int):
    """The deterministic orthogonal factor of the reference's stress smokes
    (src/bin/stress_cpu.rs:805-838): n*n draws of the 64-bit LCG
    s <- s * 6364136223846793005 + 1, x = (s >> 11) / 2^53 - 0.5, filled
    row-major, then classical Gram-Schmidt column by column against the
    already normalised columns (dot products accumulated left to right) and
    division by max(norm, 1e-300).  Returns (Q, next LCG state)."""
    vals = np.empty(n * n)
    s = state
    for k in range(n * n):
        s = (s * 6364136223846793005 + 1) & _M64
        vals[k] = (s >> 11) / float(1 << 53) - 0.5
    q = vals.reshape(n, n).copy()
    for j in range(n):
        for k in range(j):
            dot = _seq_sum(q[:, j] * q[:, k])
            q[:, j] -= dot * q[:, k]
        nrm = max(float(np.sqrt(_seq_sum(q[:, j] * q[:, j]))), 1e-300)
        q[:, j] /= nrm
    return q, s


def jade_stress_family(n: int, family: int = 6):
    """The Phase-JADE smoke family of `stress_cpu --joint`
    (src/bin/stress_cpu.rs:474-486): M_k = Q diag(1 + 0.2 k + |sin(i + 1)|) Q^T,
    k < 6, Q = stress_orthogonal(n, 0x9e3779b97f4a7c15).  The reference clamps
    n to [4, 256].  Returns (mats (family, n, n), Q)."""
    q, _ = stress_orthogonal(n, STRESS_SEED_JADE)
    i = np.arange(n, dtype=np.float64)
    mats = np.stack([(q * (1.0 + 0.2 * k + np.abs(np.sin(i + 1.0)))[None, :]) @ q.T for k in range(family)])
    return np.ascontiguousarray(mats), q


def joint_svd_stress_family(n: int, family: int = 4):
    """The two-sided family of `stress_cpu --joint-svd`
    (src/bin/stress_cpu.rs:597-612): A_k = U0 diag(1 + 0.15 k + 0.05 i) V0^T,
    k < 4, U0 then V0 drawn in turn from one stress_orthogonal stream seeded
    0xa24baed4963ee407.  The reference clamps n to [4, 128].  Returns
    (mats (family, n, n), U0, V0)."""
    u0, s = stress_orthogonal(n, STRESS_SEED_JOINT_SVD)
    v0, _ = stress_orthogonal(n, s)
    i = np.arange(n, dtype=np.float64)
    mats = np.stack([(u0 * (1.0 + 0.15 * k + 0.05 * i)[None, :]) @ v0.T for k in range(family)])
    return np.ascontiguousarray(mats), u0, v0


def polish_batch(count: int = 2048, seed: int = SEED_BATCH + 11, eps: float = 1e-2):
    """Digital-polish inputs: square A = U diag(s) V^T with n log-uniform in
    [8, 128], s log-uniform over kappa up to 1e8, and bases U0, V0 that are
    U and V perturbed by eps (Gaussian) and re-orthonormalised (QR with the
    sign of diag(R) folded in) -- what an analog or phase-flow preconditioner
    hands to the polish.  Returns (mats, U0s, Vt0s, sigmas)."""
    mats, u0s, vt0s, sig = [], [], [], []
    for b in range(count):
        s0 = rng.mix_seed(seed, b)
        n = int(np.exp(np.log(8.0) + (np.log(128.0) - np.log(8.0)) * rng.uniform(rng.mix_seed(s0, 0), 1)[0]))
        n = max(8, min(128, n))
        kap = np.exp(np.log(1e8) * rng.uniform(rng.mix_seed(s0, 1), 1)[0])
        s = np.sort(np.exp(-np.log(kap) * rng.uniform(rng.mix_seed(s0, 2), n)))[::-1]
        u = rng.haar(rng.mix_seed(s0, 3), n)
        v = rng.haar(rng.mix_seed(s0, 4), n)
        a = (u * s[None, :]) @ v.T
        gu = rng.normal(rng.mix_seed(s0, 5), n * n).reshape(n, n)
        gv = rng.normal(rng.mix_seed(s0, 6), n * n).reshape(n, n)
        qu, ru = np.linalg.qr(u + eps * gu)
        qv, rv = np.linalg.qr(v + eps * gv)
        mats.append(np.ascontiguousarray(a))
        u0s.append(np.ascontiguousarray(qu * np.sign(np.diag(ru))[None, :]))
        vt0s.append(np.ascontiguousarray((qv * np.sign(np.diag(rv))[None, :]).T))
        sig.append(s)
    return mats, u0s, vt0s, sig


def recovery_error(q, v):
    """Distance between planted Q and recovered V up to column permutation
    and sign: match columns greedily by |Q^T V| (a permutation when the
    recovery is good) and return max |1 - |<q_i, v_pi(i)>||."""
    c = np.abs(q.T @ v)
    perm = c.argmax(axis=1)
    if len(set(perm.tolist())) != len(perm):
        return 1.0
    return float(np.max(np.abs(1.0 - c[np.arange(len(perm)), perm])))


def polar_matrix(m: int, n: int, seed: int = SEED_POLAR, lo: float = 1e-6, hi: float = 1.0):
    """A = U diag(s) V^T with Haar U (m x n), V (n x n), s log-uniform in
    [lo, hi].  Returns (A, s, U, V).  The exact polar factor is U V^T."""
    s = np.exp(np.log(lo) + (np.log(hi) - np.log(lo)) * rng.uniform(rng.mix_seed(seed, 7), n))
    u = rng.haar(rng.mix_seed(seed, m * 131 + n), m, n)
    v = rng.haar(rng.mix_seed(seed, n * 977 + 3), n)
    return (u * s[None, :]) @ v.T, s, u, v
