"""Pins the oracle's Phase-BSS (LieSvdBss::separate, src/lie_svd_bss.rs:
58-101) on the reference's own test (:222-250) and on what the mathematics
fixes: the whitened channels have the identity covariance, W^T W is the
inverse covariance (independent of the singular vectors' signs), the lagged
family equals an independent numpy product, the floor of a rank-deficient
covariance, the fallback family when no lag is usable, and the unmixing /
separated consistency (no GPU)."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from lie_cliffalg_analog_svd_amd.bss import channel_phase_coherence, estimate_sir_db, synthetic_sources


def mixtures(seed, ch, S):
    rng = np.random.default_rng(seed)
    t = np.arange(S)
    src = np.stack([np.sin(2 * np.pi * (3 + 2 * i) * t / S + i) + 0.1 * rng.standard_normal(S) for i in range(ch)])
    mix = np.eye(ch) + 0.3 * rng.standard_normal((ch, ch))
    return src, mix @ src


def numpy_lagged(z, lag):
    S = z.shape[1]
    c = z[:, lag:] @ z[:, :S - lag].T / (S - lag)
    return 0.5 * (c + c.T)


def test_reference_unit_test():
    src, _mix, obs = synthetic_sources(4, 512)
    r = oracle.bss_separate(obs)
    before = estimate_sir_db(src, obs)
    after = estimate_sir_db(src, r["separated"])
    assert after > before + 3.0, (before, after)
    assert r["joint"]["final_offdiag"] <= r["joint"]["initial_offdiag"] + 1e-10
    assert np.isfinite(r["mean_channel_coherence"])


@pytest.mark.parametrize("ch,S", [(2, 64), (4, 512), (6, 300), (8, 1000)])
def test_whitening_against_numpy(ch, S):
    _src, x = mixtures(ch * 1000 + S, ch, S)
    w, z, fam, sg, fl = oracle.bss_whiten(x)
    xc = x - x.mean(axis=1, keepdims=True)
    cov = xc @ xc.T / S
    assert np.allclose(np.sort(sg)[::-1], np.linalg.eigvalsh(cov)[::-1], rtol=1e-12, atol=1e-14)
    assert fl == pytest.approx(1e-10 * max(sg.max(), 1.0))
    assert np.abs(w @ cov @ w.T - np.eye(ch)).max() < 1e-12
    assert np.abs(w.T @ w - np.linalg.inv(cov)).max() < 1e-10 * np.abs(np.linalg.inv(cov)).max()
    assert np.abs(z - w @ xc).max() < 1e-12 * np.abs(z).max()
    assert np.abs(z @ z.T / S - np.eye(ch)).max() < 1e-12
    assert fam.shape == (4, ch, ch)
    for q, lag in enumerate((1, 2, 3, 5)):
        assert np.abs(fam[q] - numpy_lagged(z, lag)).max() < 1e-13
        assert np.array_equal(fam[q], fam[q].T)


def test_unusable_lags_and_fallback_family():
    _src, x = mixtures(7, 3, 40)
    _w, z, fam, _sg, _fl = oracle.bss_whiten(x, lags=(0, 40, 55, 2))
    assert fam.shape == (1, 3, 3)  # only lag 2 is usable
    assert np.abs(fam[0] - numpy_lagged(z, 2)).max() < 1e-13
    _w, z, fam, _sg, _fl = oracle.bss_whiten(x, lags=(0, 40))
    assert fam.shape == (1, 3, 3)  # none: the whitened covariance (:192-194)
    assert np.abs(fam[0] - z @ z.T / 40).max() < 1e-13
    assert np.abs(fam[0] - np.eye(3)).max() < 1e-12


def test_rank_deficient_covariance_uses_the_floor():
    _src, x = mixtures(11, 3, 256)
    x = np.vstack([x, x[0]])  # a repeated channel: one zero covariance eigenvalue
    w, z, _fam, sg, fl = oracle.bss_whiten(x)
    assert sg.min() < fl
    assert np.all(np.isfinite(w)) and np.all(np.isfinite(z))
    # the floored direction is scaled by 1 / sqrt(floor) and not beyond
    assert np.abs(w).max() <= 1.0 / np.sqrt(fl) * (1 + 1e-12)


def test_unmix_consistency_and_coherence():
    _src, x = mixtures(5, 5, 700)
    r = oracle.bss_separate(x)
    xc = x - x.mean(axis=1, keepdims=True)
    assert np.abs(r["separated"] - r["unmixing"] @ xc).max() < 1e-11 * np.abs(r["separated"]).max()
    v = r["basis"]
    assert np.abs(v.T @ v - np.eye(5)).max() < 1e-13
    assert np.allclose(r["channel_coherence"], channel_phase_coherence(r["separated"]), rtol=1e-13, atol=1e-15)
    assert np.all((r["channel_coherence"] >= 0) & (r["channel_coherence"] <= 1))


def test_sir_metric_closed_forms():
    src, _mix, _obs = synthetic_sources(4, 512)
    perfect = estimate_sir_db(src, src[::-1] * np.array([[2.0], [-1.0], [0.5], [3.0]]))
    cross = estimate_sir_db(src, src)
    assert perfect == pytest.approx(cross)  # scale, sign and order do not matter
    rng = np.random.default_rng(3)
    q, _ = np.linalg.qr(rng.standard_normal((4, 4)))
    assert estimate_sir_db(src, q @ src) < cross
