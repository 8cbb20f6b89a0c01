"""GPU parity of the MFMA GEMMs and the Newton-Schulz polar path.

GEMM tolerances: FP64 -- 1e-13 relative to ||A|| ||B|| k^(1/2) scale against
a float64 product; FP32 -- 1e-5; BF16 operands (FP32 accumulation) --
against a float64 product of the BF16-rounded operands at 1e-5 (what is
left is accumulation order), which also pins the MFMA lane layouts.
Polar: FP64 path vs the oracle's newton_schulz_polar (same iteration) to
1e-12 and the exact U V^T of the synthetic input; mixed-precision paths
must be orthogonal to FP64 precision and within their stated error bound.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from lie_cliffalg_analog_svd_amd import (LSVD_PREC_BF16, LSVD_PREC_F32, LSVD_PREC_F64, LSVD_PTR_DEVICE, Handle,
                                         newton_schulz_polar, polar, rng, workloads)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dh():
    h = Handle(0)
    h.set_pointer_mode(LSVD_PTR_DEVICE)
    yield h
    h.close()


def _bf16(x):
    return torch.from_numpy(x.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()


@pytest.mark.parametrize("trans", [False, True])
@pytest.mark.parametrize("m,n,k", [(64, 64, 64), (100, 37, 53), (128, 256, 200), (257, 129, 65)])
def test_gemm_f64(dh, trans, m, n, k):
    print("gemm f64", trans, m, n, k, flush=True)
    a = rng.normal_matrix(1, k, m) if trans else rng.normal_matrix(1, m, k)
    b = rng.normal_matrix(2, k, n)
    ad, bd = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    cd = torch.zeros((m, n), dtype=torch.float64, device="cuda")
    polar.gemm_device(dh, trans, m, n, k, ad, bd, cd, LSVD_PREC_F64)
    dh.synchronize()
    want = (a.T if trans else a) @ b
    assert np.max(np.abs(cd.cpu().numpy() - want)) < 1e-13 * np.sqrt(k) * 8


@pytest.mark.parametrize("prec", [LSVD_PREC_F32, LSVD_PREC_BF16])
@pytest.mark.parametrize("trans", [False, True])
@pytest.mark.parametrize("m,n,k", [(128, 128, 128), (100, 37, 53), (256, 384, 160), (300, 129, 97)])
def test_gemm_lowp(dh, prec, trans, m, n, k):
    print("gemm lowp", prec, trans, m, n, k, flush=True)
    a = rng.normal_matrix(3, k, m) if trans else rng.normal_matrix(3, m, k)
    b = rng.normal_matrix(4, k, n)
    ad = torch.from_numpy(a.astype(np.float32)).cuda()
    bd = torch.from_numpy(b.astype(np.float32)).cuda()
    cd = torch.zeros((m, n), dtype=torch.float32, device="cuda")
    polar.gemm_device(dh, trans, m, n, k, ad, bd, cd, prec)
    dh.synchronize()
    aa = (a.T if trans else a)
    if prec == LSVD_PREC_BF16:
        want = _bf16(aa) @ _bf16(b)
    else:
        want = aa.astype(np.float32).astype(np.float64) @ b.astype(np.float32).astype(np.float64)
    err = np.max(np.abs(cd.cpu().numpy().astype(np.float64) - want))
    assert err < 1e-5 * np.sqrt(k) * 8, err


@pytest.mark.parametrize("m,n", [(64, 64), (96, 64), (150, 97), (256, 256)])
def test_polar_f64_vs_oracle(m, n):
    a, s, u, v = workloads.polar_matrix(m, n, seed=11, lo=1e-3)
    print("polar f64", m, n, flush=True)
    q, ok = newton_schulz_polar(a, LSVD_PREC_F64)
    qo, oko, it_o = oracle.newton_schulz_polar(a)
    info = polar.last_info
    assert ok and oko
    assert abs(info["f64_iters"] - it_o) <= 1
    assert np.linalg.norm(q - qo) / np.sqrt(n) < 1e-12
    assert np.linalg.norm(q - u @ v.T) / np.sqrt(n) < 1e-11
    assert info["orth_err"] < 1e-13 * np.sqrt(n) * 10


@pytest.mark.parametrize("m,n,lo", [(256, 256, 1e-6), (384, 128, 1e-10), (200, 200, 1.0)])
def test_polar_f64_scaled_schedule(m, n, lo):
    """LSVD_PREC_F64_SCALED reaches the exact polar factor U V^T as closely
    as the reference iteration does (the FP64 limit: rounding amplified by
    the condition number in the small-sigma directions), in fewer steps."""
    from lie_cliffalg_analog_svd_amd import LSVD_PREC_F64_SCALED
    a, s, u, v = workloads.polar_matrix(m, n, seed=13, lo=lo)
    print("polar scaled", m, n, lo, flush=True)
    q, ok = newton_schulz_polar(a, LSVD_PREC_F64_SCALED)
    info = polar.last_info
    qo, oko, it_o = oracle.newton_schulz_polar(a)
    assert ok and oko
    exact = u @ v.T
    e_scaled = np.linalg.norm(q - exact) / np.sqrt(n)
    e_ref = np.linalg.norm(qo - exact) / np.sqrt(n)
    assert e_scaled <= max(1e-13, 4.0 * e_ref), (e_scaled, e_ref)
    assert info["orth_err"] < 1e-13 * np.sqrt(n) * 10
    if lo < 1e-3:
        assert info["f64_iters"] <= 0.8 * it_o, (info["f64_iters"], it_o)


@pytest.mark.parametrize("prec", [LSVD_PREC_F32, LSVD_PREC_BF16])
def test_polar_mixed_precision_well_conditioned(prec):
    """Mixed precision at kappa = 1e2: orthogonal to FP64 precision after the
    FP64 refinement; the error against the exact polar factor is that of the
    low precision (u_low kappa), and the A-dependent residual
    skew_err = ||Q^T A - A^T Q||_F / ||A||_F sees it, so the factor is only
    trusted where that residual is below 1e-8 sqrt(n)."""
    m, n = 384, 256
    a, s, u, v = workloads.polar_matrix(m, n, seed=12, lo=1e-2)
    print("polar mixed", prec, flush=True)
    q, ok = newton_schulz_polar(a, prec)
    info = polar.last_info
    print("  info", info, flush=True)
    assert info["orth_err"] < 1e-12 * np.sqrt(n)
    assert info["lowp_iters"] > 0 and info["f64_iters"] <= 8
    err = np.linalg.norm(q - u @ v.T, 2)
    skew_exact = np.linalg.norm(q.T @ a - a.T @ q) / np.linalg.norm(a)
    assert abs(info["skew_err"] - skew_exact) <= 1e-12 + 1e-6 * skew_exact
    if prec == LSVD_PREC_F32:
        assert err < 50 * 2.0 ** -24 * 1e2, err
    else:
        assert 1e-4 < err < 0.5, err
        assert not ok and info["skew_err"] > 1e-8 * np.sqrt(n)
    assert ok == (info["skew_err"] <= 1e-8 * np.sqrt(n))


@pytest.mark.parametrize("m,n", [(512, 512), (512, 256)])
def test_polar_kappa_1e6_vs_oracle(m, n):
    """Config-3 spectra (s log-uniform in [1e-6, 1]) at a size the oracle
    runs in seconds.  FP64 (plain and scaled steps) equals the oracle's
    newton_schulz_polar to the FP64 limit of this conditioning and is
    trusted; the FP32 and BF16 iterations end orthogonal to FP64 precision
    but rotated away from A's polar factor in its small-sigma directions
    (u_low kappa > 1): the skew residual flags them (polar_ok = 0) and their
    error against the oracle is stated, not hidden."""
    from lie_cliffalg_analog_svd_amd import LSVD_PREC_F64_SCALED
    a, s, u, v = workloads.polar_matrix(m, n, seed=0x5EED0003, lo=1e-6)
    print("polar kappa 1e6", m, n, flush=True)
    qo, oko, it_o = oracle.newton_schulz_polar(a)
    assert oko
    rootn = np.sqrt(n)
    for prec in (LSVD_PREC_F64, LSVD_PREC_F64_SCALED):
        q, ok = newton_schulz_polar(a, prec)
        info = polar.last_info
        print("  f64", prec, info, flush=True)
        assert ok and info["skew_err"] < 1e-12
        assert np.linalg.norm(q - qo) / rootn < 1e-9
    for prec in (LSVD_PREC_F32, LSVD_PREC_BF16):
        q, ok = newton_schulz_polar(a, prec)
        info = polar.last_info
        err = np.linalg.norm(q - qo) / rootn
        print("  lowp", prec, info, "err vs oracle", err, flush=True)
        assert info["orth_err"] < 1e-12 * rootn
        assert not ok and info["skew_err"] > 1e-8 * rootn
        assert err > 1e-6


@pytest.mark.parametrize("m,n", [(4096, 4096), (8192, 2048)])
def test_polar_config3_full_size(m, n):
    """Config 3 at full size: FP64 scaled steps agree with the plain FP64
    iteration to 1e-9 (Frobenius / sqrt(n)) and both are trusted; the
    low-precision routes are orthogonal to FP64 precision and report
    polar_ok = 0 through the skew residual, with their error against the
    FP64 factor stated (measured, see the bench's polar leg)."""
    import torch
    from lie_cliffalg_analog_svd_amd import LSVD_PREC_F64_SCALED
    a, s, u, v = workloads.polar_matrix(m, n, seed=0x5EED0003, lo=1e-6)
    print("polar config 3", m, n, flush=True)
    h = Handle(0)
    h.set_pointer_mode(LSVD_PTR_DEVICE)
    ad = torch.from_numpy(a).cuda()
    out = {}
    for prec in (LSVD_PREC_F64, LSVD_PREC_F64_SCALED, LSVD_PREC_F32, LSVD_PREC_BF16):
        qd = torch.empty_like(ad)
        info = polar.polar_device(h, ad, m, n, qd, prec)
        h.synchronize()
        out[prec] = (qd, info)
        print("  ", prec, info, flush=True)
    h.close()
    rootn = np.sqrt(n)
    q64 = out[LSVD_PREC_F64][0]
    for prec, (qd, info) in out.items():
        assert info["orth_err"] < 1e-12 * rootn
        e = float(torch.linalg.norm(qd - q64)) / rootn
        if prec in (LSVD_PREC_F64, LSVD_PREC_F64_SCALED):
            assert info["polar_ok"] == 1 and info["skew_err"] < 1e-12 and e < 1e-9, (prec, e)
        else:
            assert info["polar_ok"] == 0 and info["skew_err"] > 1e-8 * rootn and e > 1e-6, (prec, e)
