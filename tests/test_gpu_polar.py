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
def test_polar_mixed_precision(prec):
    """Mixed precision: orthogonal to FP64 precision; error against the
    exact polar factor bounded by the low-precision unit roundoff times
    the condition number (kappa = 1e2 here)."""
    m, n = 384, 256
    a, s, u, v = workloads.polar_matrix(m, n, seed=12, lo=1e-2)
    print("polar mixed", prec, flush=True)
    q, ok = newton_schulz_polar(a, prec)
    info = polar.last_info
    assert ok
    assert info["orth_err"] < 1e-12 * np.sqrt(n)
    assert info["lowp_iters"] > 0 and info["f64_iters"] <= 8
    err = np.linalg.norm(q - u @ v.T, 2)
    unit = 2.0 ** -8 if prec == LSVD_PREC_BF16 else 2.0 ** -24
    assert err < 50 * unit * 1e2, err
