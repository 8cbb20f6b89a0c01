"""Phase-BSS (src/lie_svd_bss.rs) on the GPU.

``LieSvdBss.separate`` whitens a channels x samples observation block with
LieSvdSmall, joint-diagonalizes the lagged covariances of the whitened
channels with Phase-JADE and returns the separated channels.
``separate_batched`` takes many blocks of one shape at once: their
covariances go through one batched LieSvdSmall call and their lagged families
through one batched Phase-JADE call.  ``channel_phase_coherence`` and
``estimate_sir_db`` are the reference's host-side diagnostics.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .handle import Handle, _ptr, default_handle
from .joint import JointDiagonalizationParams


@dataclass
class PhaseBssParams:
    """src/lie_svd_bss.rs:13-33."""
    lags: list = field(default_factory=lambda: [1, 2, 3, 5])
    eps: float = 1e-10
    joint: JointDiagonalizationParams = field(default_factory=lambda: JointDiagonalizationParams.for_n(4))

    @staticmethod
    def for_channels(channels: int) -> "PhaseBssParams":
        return PhaseBssParams([1, 2, 3, 5], 1e-10, JointDiagonalizationParams.for_n(channels))

    @staticmethod
    def default() -> "PhaseBssParams":
        return PhaseBssParams.for_channels(4)


@dataclass
class PhaseBssTrace:
    """src/lie_svd_bss.rs:36-44."""
    channels: int
    samples: int
    covariance_floor: float
    joint_initial_offdiag: float
    joint_final_offdiag: float
    joint_rotations: int
    mean_channel_coherence: float


@dataclass
class PhaseBssResult:
    """src/lie_svd_bss.rs:47-52 (plus the whitening, for diagnostics)."""
    unmixing: np.ndarray
    separated: np.ndarray
    channel_coherence: np.ndarray
    trace: PhaseBssTrace
    whitening: np.ndarray | None = None


def separate_batched(observations, params: PhaseBssParams | None = None, handle: Handle | None = None,
                     with_whitening: bool = False):
    """observations: (B, channels, samples) -> list of PhaseBssResult."""
    x = np.ascontiguousarray(np.asarray(observations, dtype=np.float64))
    if x.ndim != 3:
        raise _lib.ShapeError(_lib.LSVD_E_SHAPE, "PhaseBSS: observations must be (B, channels, samples)")
    B, ch, S = x.shape
    params = params or PhaseBssParams.for_channels(ch)
    h = handle or default_handle()
    lags = np.ascontiguousarray(np.asarray(params.lags, dtype=np.int32))
    un = np.zeros((B, ch, ch)); sep = np.zeros((B, ch, S)); coh = np.zeros((B, ch))
    w = np.zeros((B, ch, ch)) if with_whitening else None
    tr = (_lib.LsvdBssTrace * max(B, 1))()
    cp = params.joint._c()
    st = _lib.lib().lsvd_bss_separate_f64(h.raw, B, ch, S, _ptr(x), len(lags), _ptr(lags), float(params.eps),
                                          C.byref(cp), _ptr(un), _ptr(sep), _ptr(coh), _ptr(w), None, tr)
    h.check(st)
    return [PhaseBssResult(un[b], sep[b], coh[b], PhaseBssTrace(**tr[b].as_dict()),
                           None if w is None else w[b]) for b in range(B)]


class LieSvdBss:
    """src/lie_svd_bss.rs:54-102."""

    @staticmethod
    def separate(observations, params: PhaseBssParams | None = None, handle: Handle | None = None) -> PhaseBssResult:
        x = np.asarray(observations, dtype=np.float64)
        if x.ndim != 2 or x.shape[0] == 0 or x.shape[1] <= 1:
            raise _lib.ShapeError(_lib.LSVD_E_SHAPE, "PhaseBSS: empty observations")
        return separate_batched(x[None], params, handle)[0]


def channel_phase_coherence(channels) -> np.ndarray:
    """src/lie_svd_bss.rs:104-126 (host)."""
    y = np.asarray(channels, dtype=np.float64)
    n, S = y.shape
    if S <= 1:
        return np.zeros(n)
    nxt = np.roll(y, -1, axis=1)
    norm = (y * y).sum(axis=1)
    lock = np.abs((y * nxt).sum(axis=1)) / np.maximum(norm, 1e-300)
    smooth = 1.0 / (1.0 + np.sqrt(((nxt - y) ** 2).sum(axis=1) / np.maximum(norm, 1e-300)))
    return np.clip(0.5 * lock + 0.5 * smooth, 0.0, 1.0)


def _abs_corr(a, b) -> np.ndarray:
    """absolute_correlation_matrix (src/lie_svd_bss.rs:199-215)."""
    n = min(a.shape[0], b.shape[0]); S = min(a.shape[1], b.shape[1])
    a = a[:n, :S]; b = b[:n, :S]
    na = np.sqrt((a * a).sum(axis=1)); nb = np.sqrt((b * b).sum(axis=1))
    return np.abs(a @ b.T) / np.maximum(np.outer(na, nb), 1e-300)


def estimate_sir_db(reference, separated) -> float:
    """src/lie_svd_bss.rs:128-160: greedy best-correlation matching, signal
    over interference in dB."""
    r = np.asarray(reference, dtype=np.float64); s = np.asarray(separated, dtype=np.float64)
    n = min(r.shape[0], s.shape[0])
    if n == 0:
        return 0.0
    corr = _abs_corr(r, s)
    used_r = [False] * n; used_s = [False] * n
    signal = 0.0; interference = 0.0
    for _ in range(n):
        best = (0, 0, -1.0)
        for i in range(n):
            if used_r[i]:
                continue
            for j in range(n):
                if not used_s[j] and corr[i, j] > best[2]:
                    best = (i, j, corr[i, j])
        used_r[best[0]] = True; used_s[best[1]] = True
        signal += best[2] * best[2]
        for i in range(n):
            if i != best[0]:
                interference += corr[i, best[1]] ** 2
    return float(10.0 * np.log10(signal / max(interference, 1e-12)))


def synthetic_sources(channels: int = 4, samples: int = 512):
    """The reference test's sources and mixing (src/lie_svd_bss.rs:222-247):
    returns (sources, mixing, observations = mixing @ sources)."""
    t = np.arange(samples)
    x = t / samples
    rows = []
    for i in range(channels):
        if i == 0:
            rows.append(np.sin(2 * np.pi * 7 * x))
        elif i == 1:
            rows.append(np.sign(np.cos(2 * np.pi * 11 * x)))
        elif i == 2:
            rows.append(np.sin(2 * np.pi * 3 * x) + 0.4 * np.sin(2 * np.pi * 17 * x))
        else:
            rows.append(np.sin((t * 37 + 11).astype(np.float64)) * 0.7)
    src = np.stack(rows)
    mix = np.array([[1.0 if i == j else 0.25 * np.sin(i * 13 + j * 7 + 3) for j in range(channels)]
                    for i in range(channels)])
    return src, mix, mix @ src
