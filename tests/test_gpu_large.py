"""LieSvdSmall::solve of one large square matrix on the chip-wide route
(src/lie_svd_small.rs:296-323 with every step spread over the GPU: the
polar GEMMs, P = sym(Q^T A), cyclic Jacobi over the round-robin pair layers
with one matrix per family, the stable descending sort and U = Q V) against
the CPU oracle, against the workgroup-per-matrix kernels, and at the sizes
the oracle is slow on through LAPACK's singular values and the residuals.

Tolerances as for the other square solves: sigma two-tier (relative 1e-12
above 1e-3 sigma_max, absolute 64 n eps sigma_max), singular vectors per
cluster by principal angle, residuals at the FP64 level scaled by n / 32.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np
import pytest

import oracle
from lie_cliffalg_analog_svd_amd import Handle, LieSvdSmall, _lib, rng
from tests._util import assert_factors, assert_sigma_close, imposed_spectrum, uniform_matrix
from tests._util import vector_parity

pytestmark = pytest.mark.gpu

LARGE_PATH = 3


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _handle(min_n: int) -> Handle:
    """A handle whose chip-wide route starts at min_n (LSVD_LARGE_MIN_N is
    read when the handle is created)."""
    old = os.environ.get("LSVD_LARGE_MIN_N")
    os.environ["LSVD_LARGE_MIN_N"] = str(min_n)
    try:
        return Handle()
    finally:
        if old is None:
            del os.environ["LSVD_LARGE_MIN_N"]
        else:
            os.environ["LSVD_LARGE_MIN_N"] = old


def _vs_oracle(a, what, h=None):
    n = a.shape[0]
    _say("case", what, a.shape)
    u, s, vt = LieSvdSmall.solve(a, h)
    info = dict(LieSvdSmall.last_info)
    uo, so, vto, _ = oracle.small_solve(a)
    assert_factors(a, u, s, vt, 1e-11 * max(1.0, n / 32), what)
    assert_sigma_close(s, so, n, what=what)
    vector_parity(u, uo, so, n, what + " U")
    vector_parity(vt.T, vto.T, so, n, what + " V")
    return info


@pytest.mark.parametrize("n", [192, 255, 256, 320, 384])
def test_large_route_uniform_vs_oracle(n):
    info = _vs_oracle(uniform_matrix(700 + n, n, n), f"large uniform n={n}")
    assert info["path"] == LARGE_PATH and info["polar_ok"] == 1, info
    assert 3 <= info["jacobi_sweeps"] <= 20 and info["rotations"] > 0


def test_large_route_matches_workgroup_kernels():
    """The same matrix through both routes of the single-matrix API."""
    n = 224
    a = rng.normal_matrix(0x5EED0300, n, n)
    h_large, h_wg = _handle(2), _handle(100000)
    u1, s1, vt1 = LieSvdSmall.solve(a, h_large)
    assert LieSvdSmall.last_info["path"] == LARGE_PATH
    u2, s2, vt2 = LieSvdSmall.solve(a, h_wg)
    assert LieSvdSmall.last_info["path"] == 0
    assert_sigma_close(s1, s2, n, what="large vs workgroup")
    vector_parity(u1, u2, s2, n, "large vs workgroup U")
    vector_parity(vt1.T, vt2.T, s2, n, "large vs workgroup V")
    # small n on the chip-wide route (forced) against the oracle
    for m in (2, 3, 17, 64):
        info = _vs_oracle(uniform_matrix(800 + m, m, m), f"forced large n={m}", h_large)
        assert info["path"] == LARGE_PATH
    h_large.close()
    h_wg.close()


@pytest.mark.parametrize("kind", ["clusters", "kappa1e10", "graded"])
def test_large_route_imposed_spectra(kind):
    """Imposed spectra at n = 256: repeated singular values (per-cluster
    subspaces), condition 1e10, and a graded spectrum; sigma against both
    the oracle and the imposed values."""
    n = 256
    if kind == "clusters":
        sig = np.repeat([8.0, 4.0, 2.0, 1.0], n // 4)
    elif kind == "kappa1e10":
        sig = np.logspace(0, -10, n)
    else:
        sig = 1.0 + np.arange(n)[::-1] * 0.01
    a = imposed_spectrum(0x5EED0400 + len(kind), n, n, sig)
    info = _vs_oracle(a, f"large {kind}")
    assert info["path"] == LARGE_PATH
    s = LieSvdSmall.solve(a)[1]
    assert_sigma_close(s, np.sort(sig)[::-1], n, what=kind + " vs imposed")


def test_large_route_rank_deficient_takes_repair_branch():
    """A rank-150 matrix of order 200: Newton-Schulz cannot make Q
    orthogonal on the null space, so the polar factor is not trusted and the
    solve takes the reference's repair branch (U from A V / sigma, then
    Gram-Schmidt) on the workgroup-per-matrix kernels."""
    n, r = 200, 150
    sig = np.concatenate([np.linspace(2.0, 1.0, r), np.zeros(n - r)])
    a = imposed_spectrum(0x5EED0500, n, n, sig)
    _say("case rank deficient", n, r)
    u, s, vt = LieSvdSmall.solve(a)
    info = dict(LieSvdSmall.last_info)
    uo, so, vto, oinfo = oracle.small_solve(a)
    assert info["path"] == 0 and info["polar_ok"] == oinfo["polar_ok"] == 0, (info, oinfo)
    assert_factors(a, u, s, vt, 1e-10, "rank deficient")
    assert_sigma_close(s, so, n, what="rank deficient")


@pytest.mark.parametrize("n", [512, 1024])
def test_large_route_full_sizes(n):
    """The reference times N = 512 at ~7 s and N = 1024 at ~77 s
    (src/lie_svd_small.rs:51-54); here against LAPACK's singular values
    and the residual checks (the oracle takes minutes at 1024)."""
    a = rng.normal_matrix(0x5EED0600 + n, n, n)
    _say("case large full", n)
    u, s, vt = LieSvdSmall.solve(a)
    info = dict(LieSvdSmall.last_info)
    assert info["path"] == LARGE_PATH, info
    assert_factors(a, u, s, vt, 1e-11 * n / 32, f"n={n}")
    ref = np.linalg.svd(a, compute_uv=False)
    # LAPACK's own error on the smallest values is ~ kappa eps, so the
    # relative tier starts at 1e-2 sigma_max here
    assert_sigma_close(s, ref, n, rel_tol=1e-12, rel_floor=1e-2, what=f"n={n} vs LAPACK")


def test_large_route_device_pointers_match_host_mode():
    """Device pointer mode (torch buffers) gives the host mode's factors.
    Compared to rounding: the iteration's stop-test norms are workgroup
    partials added with atomics, so near a threshold two runs can stop one
    step apart."""
    import torch
    n = 256
    a = uniform_matrix(0x5EED0700, n, n)
    u1, s1, vt1 = LieSvdSmall.solve(a)
    h = Handle()
    h.set_pointer_mode(_lib.LSVD_PTR_DEVICE)
    dev = torch.device("cuda:0")
    da = torch.from_numpy(a).to(dev)
    du = torch.zeros((n, n), dtype=torch.float64, device=dev)
    ds = torch.zeros(n, dtype=torch.float64, device=dev)
    dvt = torch.zeros((n, n), dtype=torch.float64, device=dev)
    dinfo = torch.zeros(8, dtype=torch.int32, device=dev)
    p = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
    torch.cuda.synchronize()
    h.check(_lib.lib().lsvd_small_solve_f64(h.raw, p(da), n, p(du), p(ds), p(dvt), p(dinfo)))
    torch.cuda.synchronize()
    assert int(dinfo[6].item()) == LARGE_PATH
    s2 = ds.cpu().numpy()
    assert_sigma_close(s2, s1, n, what="device vs host mode")
    vector_parity(du.cpu().numpy(), u1, s1, n, "device vs host mode U")
    vector_parity(dvt.cpu().numpy().T, vt1.T, s1, n, "device vs host mode V")
    h.close()


@pytest.mark.parametrize("n", [160, 300])
def test_large_eigh_vs_oracle(n):
    """eigh_jacobi_full (src/lie_svd_small.rs:212-285) above 128 on the block
    Jacobi: eigenvalues to 64 n eps of the largest against the oracle's, the
    eigen-residual and orthogonality, per-cluster eigenvector parity."""
    from lie_cliffalg_analog_svd_amd import eigh_jacobi_full
    _say("case large eigh", n)
    g = rng.normal_matrix(0x5EED0800 + n, n, n)
    p = g + g.T
    v, e = eigh_jacobi_full(p)
    info = dict(eigh_jacobi_full.last_info)
    assert info["path"] == LARGE_PATH and info["jacobi_sweeps"] >= 3, info
    vo, eo, _ = oracle.eigh_jacobi_full(p)
    eps = np.finfo(float).eps
    assert np.all(np.diff(e) <= 0)
    assert np.max(np.abs(e - eo)) < 64 * n * eps * np.abs(eo).max()
    assert np.linalg.norm(v.T @ v - np.eye(n)) < 1e-11 * n / 32
    assert np.linalg.norm(v.T @ p @ v - np.diag(e)) < 1e-11 * n / 32 * np.linalg.norm(p)
    # eigenvectors up to sign (distinct eigenvalues of a random symmetric matrix)
    c = np.abs(np.sum(v * vo, axis=0))
    gaps = np.minimum(np.abs(np.diff(eo, prepend=np.inf)), np.abs(np.diff(eo, append=-np.inf)))
    ok = gaps > 1e-6 * np.abs(eo).max()
    assert np.max(np.abs(1.0 - c[ok])) < 1e-9


def test_large_eigh_psd_and_degenerate():
    """A PSD matrix with repeated eigenvalues (clusters of 32) at n = 256:
    eigenvalues against the imposed ones, the invariant subspaces by angle,
    and the same answer from the workgroup kernels."""
    from lie_cliffalg_analog_svd_amd import eigh_jacobi_full
    n = 256
    lam = np.repeat(np.linspace(5.0, 0.5, n // 32), 32)
    q = rng.haar(0x5EED0900, n)
    p = (q * lam[None, :]) @ q.T
    p = 0.5 * (p + p.T)
    v, e = eigh_jacobi_full(p)
    assert eigh_jacobi_full.last_info["path"] == LARGE_PATH
    assert np.max(np.abs(e - np.sort(lam)[::-1])) < 1e-12 * 5.0 * n
    h_wg = _handle(100000)
    v2, e2 = eigh_jacobi_full(p, h_wg)
    assert eigh_jacobi_full.last_info["path"] == 0
    assert np.max(np.abs(e - e2)) < 64 * n * np.finfo(float).eps * 5.0
    vector_parity(v, v2, e2, n, "eigh clusters")
    h_wg.close()


@pytest.mark.parametrize("m,n", [(400, 200), (1500, 300), (200, 450), (129, 1000)])
def test_large_rectangular_vs_oracle(m, n):
    """solve_rectangular (src/lie_svd_small.rs:340-360) with min(m, n) > 128
    on the chip-wide route: a tall matrix runs the polar route on itself (the
    reference reduces it by MGS QR first; both give the thin SVD), a wide one
    is solved transposed.  Oracle parity with the two-tier sigma bar and
    per-cluster subspaces, thin shapes, residuals."""
    a = uniform_matrix(0x5EED0A00 + m + n, m, n)
    _say("case large rect", m, n)
    u, s, vt = LieSvdSmall.solve_rectangular(a)
    info = dict(LieSvdSmall.last_info)
    k = min(m, n)
    assert info["path"] == LARGE_PATH, info
    assert u.shape == (m, k) and s.shape == (k,) and vt.shape == (k, n)
    uo, so, vto, _ = oracle.small_solve_rectangular(a)
    assert_factors(a, u, s, vt, 1e-11 * max(1.0, max(m, n) / 32), f"rect {m}x{n}")
    assert_sigma_close(s, so, max(m, n), what=f"rect {m}x{n}")
    vector_parity(u, uo, so, k, f"rect {m}x{n} U")
    vector_parity(vt.T, vto.T, so, k, f"rect {m}x{n} V")


def test_large_rectangular_rank_deficient_takes_workgroup_path():
    """A tall 300 x 150 matrix of rank 100: the polar factor of A itself is
    not orthogonal on A's null space, so the solve goes to the workgroup
    kernels (the reference's MGS QR front end and repair)."""
    m, n, r = 300, 150, 100
    rng_ = np.random.default_rng(0x5EED0B00)
    a = rng_.standard_normal((m, r)) @ rng_.standard_normal((r, n))
    _say("case large rect rank deficient", m, n, r)
    u, s, vt = LieSvdSmall.solve_rectangular(a)
    assert LieSvdSmall.last_info["path"] == 0
    uo, so, vto, _ = oracle.small_solve_rectangular(a)
    assert_sigma_close(s, so, m, what="rect rank deficient")
    m_ = np.linalg.norm(u * s[None, :] @ vt - a) / np.linalg.norm(a)
    assert m_ < 1e-11


def test_large_route_nonfinite_and_zero_inputs():
    """Inputs the polar step cannot trust take the workgroup kernels' path
    (path 0), as the reference's Gram-Schmidt fallback and repair do: a NaN
    entry (the polar iteration flags the non-finite iterate) and the zero
    matrix (X stays 0, so ||Q^T Q - I|| = sqrt(n)).  The eigh route hands a
    NaN diagonal to the workgroup kernels the same way."""
    from lie_cliffalg_analog_svd_amd import eigh_jacobi_full
    n = 200
    _say("case large nonfinite/zero", n)
    a = uniform_matrix(0x5EED0C00, n, n)
    a[3, 7] = np.nan
    u, s, vt = LieSvdSmall.solve(a)
    assert LieSvdSmall.last_info["path"] == 0
    z = np.zeros((n, n))
    u, s, vt = LieSvdSmall.solve(z)
    assert LieSvdSmall.last_info["path"] == 0 and np.all(s == 0)
    assert np.all(np.isfinite(u)) and np.all(np.isfinite(vt))
    p = a + a.T
    v, e = eigh_jacobi_full(p)
    assert eigh_jacobi_full.last_info["path"] == 0
