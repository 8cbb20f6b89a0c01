"""GPU parity of Phase-BSS (LieSvdBss::separate, src/lie_svd_bss.rs:58-101)
against the oracle: the reference's test, whitening, unmixing and separated
channels on mixtures of several shapes, batches against single calls, the
floor and fallback paths, device pointers and the error paths.

The covariance and lagged sums run in another order than the oracle's (split
over the samples, summed per split), so everything agrees to rounding; the
whitening is compared through W C W^T = I and W^T W = C^-1, which do not
depend on the signs LieSvdSmall gives the singular vectors, and the
separated channels / unmixing rows up to sign.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from lie_cliffalg_analog_svd_amd import (LSVD_PTR_DEVICE, Handle, LieSvdBss, LsvdError, PhaseBssParams, ShapeError,
                                         estimate_sir_db, separate_batched)
from lie_cliffalg_analog_svd_amd import _lib
from lie_cliffalg_analog_svd_amd.bss import synthetic_sources
from tests.test_oracle_bss import mixtures

pytestmark = pytest.mark.gpu


def rows_up_to_sign(a, b):
    s = np.sign(np.sum(a * b, axis=1, keepdims=True))
    s[s == 0] = 1.0
    return np.abs(a * s - b).max()


def check_vs_oracle(x, r, what, tol=1e-9):
    o = oracle.bss_separate(x)
    ch, S = x.shape
    xc = x - x.mean(axis=1, keepdims=True)
    cov = xc @ xc.T / S
    w = r.whitening
    if w is not None:
        # whitening from an FP64 eigendecomposition: first-order error
        # ch eps kappa(cov) (kappa ~ 2e4 at 64 channels of 300 samples)
        w_tol = max(1e-11, ch * np.finfo(float).eps * np.linalg.cond(cov))
        assert np.abs(w @ cov @ w.T - np.eye(ch)).max() < w_tol, what
        inv = np.linalg.inv(cov)
        assert np.abs(w.T @ w - inv).max() < 1e-10 * np.abs(inv).max(), what
    assert rows_up_to_sign(r.unmixing, o["unmixing"]) < tol * np.abs(o["unmixing"]).max(), what
    assert rows_up_to_sign(r.separated, o["separated"]) < tol * np.abs(o["separated"]).max(), what
    assert np.allclose(r.channel_coherence, o["channel_coherence"], rtol=0, atol=1e-10), what
    assert r.trace.covariance_floor == pytest.approx(o["covariance_floor"], rel=1e-12), what
    assert r.trace.joint_initial_offdiag == pytest.approx(o["joint"]["initial_offdiag"], rel=1e-10), what
    assert r.trace.joint_final_offdiag <= r.trace.joint_initial_offdiag + 1e-10, what
    assert r.trace.channels == ch and r.trace.samples == S
    assert np.abs(r.separated - r.unmixing @ xc).max() < 1e-11 * np.abs(r.separated).max(), what


def test_reference_unit_test():
    src, _mix, obs = synthetic_sources(4, 512)
    r = LieSvdBss.separate(obs, PhaseBssParams.for_channels(4))
    before = estimate_sir_db(src, obs)
    after = estimate_sir_db(src, r.separated)
    assert after > before + 3.0, (before, after)
    assert r.trace.joint_final_offdiag <= r.trace.joint_initial_offdiag + 1e-10
    assert np.isfinite(r.trace.mean_channel_coherence)
    o = oracle.bss_separate(obs)
    assert after == pytest.approx(estimate_sir_db(src, o["separated"]), abs=1e-6)
    check_vs_oracle(obs, separate_batched(obs[None], with_whitening=True)[0], "reference test")


@pytest.mark.parametrize("ch,S", [(2, 64), (3, 8), (4, 513), (8, 1000), (16, 4096), (64, 300)])
def test_mixtures_vs_oracle(ch, S):
    _src, x = mixtures(ch * 7919 + S, ch, S)
    print(f"bss ch={ch} S={S}", flush=True)
    r = separate_batched(x[None], with_whitening=True)[0]
    check_vs_oracle(x, r, f"ch={ch} S={S}")


def test_batch_matches_single_calls():
    xs = np.stack([mixtures(100 + b, 6, 800)[1] for b in range(5)])
    rb = separate_batched(xs, with_whitening=True)
    for b in range(5):
        rs = LieSvdBss.separate(xs[b])
        assert rows_up_to_sign(rb[b].separated, rs.separated) < 1e-10 * np.abs(rs.separated).max()
        assert rows_up_to_sign(rb[b].unmixing, rs.unmixing) < 1e-10 * np.abs(rs.unmixing).max()
        check_vs_oracle(xs[b], rb[b], f"batch item {b}")


def test_floor_and_fallback_family():
    _src, x = mixtures(11, 3, 256)
    x = np.vstack([x, x[0]])  # rank-deficient covariance: the floor applies
    r = separate_batched(x[None], with_whitening=True)[0]
    _w, _z, _fam, sg, fl = oracle.bss_whiten(x)
    assert r.trace.covariance_floor == pytest.approx(fl, rel=1e-12)
    assert np.all(np.isfinite(r.separated)) and np.all(np.isfinite(r.unmixing))
    assert np.abs(r.whitening).max() <= 1.0 / np.sqrt(fl) * (1 + 1e-9)
    # no usable lag: the family is the whitened covariance (~I), nothing to rotate
    _src, y = mixtures(12, 3, 40)
    p = PhaseBssParams.for_channels(3)
    p.lags = [0, 40, 77]
    r = separate_batched(y[None], p, with_whitening=True)[0]
    w, z, fam, _sg, _fl = oracle.bss_whiten(y, lags=(0, 40, 77))
    assert fam.shape == (1, 3, 3)
    assert np.abs(r.separated @ r.separated.T / 40 - np.eye(3)).max() < 1e-11
    # two samples (the smallest the reference accepts): rank-one covariance
    _src, y2 = mixtures(13, 3, 2)
    r = separate_batched(y2[None])[0]
    assert np.all(np.isfinite(r.separated)) and np.all(np.isfinite(r.unmixing))
    yc = y2 - y2.mean(axis=1, keepdims=True)
    assert np.abs(r.separated - r.unmixing @ yc).max() < 1e-9 * max(np.abs(r.separated).max(), 1.0)


def test_device_pointers():
    xs = np.stack([mixtures(200 + b, 5, 600)[1] for b in range(3)])
    ref = separate_batched(xs)
    h = Handle()
    h.set_pointer_mode(LSVD_PTR_DEVICE)
    B, ch, S = xs.shape
    dx = torch.from_numpy(xs).cuda()
    un = torch.zeros((B, ch, ch), dtype=torch.float64, device="cuda")
    sep = torch.zeros((B, ch, S), dtype=torch.float64, device="cuda")
    coh = torch.zeros((B, ch), dtype=torch.float64, device="cuda")
    lags = np.array([1, 2, 3, 5], dtype=np.int32)
    cp = PhaseBssParams.for_channels(ch).joint._c()
    import ctypes as C
    tr = (_lib.LsvdBssTrace * B)()
    st = _lib.lib().lsvd_bss_separate_f64(h.raw, B, ch, S, C.c_void_p(dx.data_ptr()), 4,
                                          lags.ctypes.data_as(C.c_void_p), 1e-10, C.byref(cp),
                                          C.c_void_p(un.data_ptr()), C.c_void_p(sep.data_ptr()),
                                          C.c_void_p(coh.data_ptr()), None, None, tr)
    h.check(st)
    torch.cuda.synchronize()
    for b in range(B):
        assert np.abs(sep[b].cpu().numpy() - ref[b].separated).max() < 1e-12 * np.abs(ref[b].separated).max()
        assert np.abs(un[b].cpu().numpy() - ref[b].unmixing).max() < 1e-12 * np.abs(ref[b].unmixing).max()
        assert tr[b].joint_rotations == ref[b].trace.joint_rotations


def test_error_paths():
    with pytest.raises(ShapeError):
        LieSvdBss.separate(np.zeros((3, 1)))
    with pytest.raises(LsvdError):
        separate_batched(np.zeros((1, 65, 16)))
    assert separate_batched(np.zeros((0, 4, 16))) == []
