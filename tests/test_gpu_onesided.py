"""GPU parity of the QR + one-sided Jacobi batch route
(LSVD_PREC_F64_ONESIDED) against the CPU oracle, through the C ABI.

The route factors A = Q R (Householder) and runs one-sided Jacobi on R^T
(the same thin SVD as LieSvdSmall::solve_rectangular, src/lie_svd_small.rs:
340-360, by another FP64 method), so its outputs are held to the same bar
as the reference-iteration kernels: singular values two-tier (relative
<= 1e-12 for sigma_i >= 1e-3 sigma_max, absolute <= 64 n eps sigma_max),
singular vectors per cluster of equal singular values within
1e3 n eps sigma_max / gap of the oracle's, residuals at the FP64 level.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle
from lie_cliffalg_analog_svd_amd import LSVD_PREC_F64_ONESIDED, LSVD_PREC_F64_SCALED, rng, svd_batched, workloads
from tests._util import EPS, assert_factors, assert_sigma_close, imposed_spectrum, uniform_matrix
from tests._util import vector_parity

pytestmark = pytest.mark.gpu

OS = LSVD_PREC_F64_ONESIDED
_HANDLES = {}


def _handle_env(**env):
    """A handle created under the given environment (read at creation)."""
    import os
    from lie_cliffalg_analog_svd_amd import Handle
    old = {k: os.environ.get(k) for k in env}
    os.environ.update({k: str(v) for k, v in env.items()})
    try:
        return Handle(0)
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


def _handle():
    """A handle that sends every padded size (16..128) to the one-sided
    kernels."""
    if "os" not in _HANDLES:
        _HANDLES["os"] = _handle_env(LSVD_OS_MAX_NP=128)
    return _HANDLES["os"]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _check_batch(mats, what, refs=None, resid=1e-11):
    _say("case", what, len(mats), "matrices")
    out, info = svd_batched(mats, precision=OS, handle=_handle(), with_info=True)
    ref = oracle.svd_batch(mats)
    for b, (a, (u, s, vt), (uo, so, vto)) in enumerate(zip(mats, out, ref)):
        n = max(a.shape)
        tag = f"{what}[{b}] {a.shape}"
        assert_factors(a, u, s, vt, resid * max(1.0, n / 32), tag)
        oracle_ok = True
        if refs is not None and refs[b] is not None:
            # the imposed spectrum is the truth; the oracle's MGS front end
            # (qr_reduce, as in the reference) loses ~eps kappa sigma_max on
            # tall ill-conditioned inputs, so against the oracle the bound is
            # widened by the oracle's own error
            truth = np.sort(refs[b])[::-1]
            assert_sigma_close(s, truth, n, what=tag + " vs imposed")
            o_err = float(np.max(np.abs(so - truth)) / truth[0])
            g_err = float(np.max(np.abs(s - so)) / truth[0])
            oracle_ok = o_err <= 64 * n * EPS
            assert g_err <= max(64 * n * EPS, 2.0 * o_err), tag
        if oracle_ok:
            assert_sigma_close(s, so, n, what=tag + " vs oracle")
            vector_parity(u, uo, so, n, tag + " U")
            vector_parity(vt.T, vto.T, so, n, tag + " V")
        if not (a.shape[0] == a.shape[1] and a.shape[0] <= 4):
            assert info[b]["path"] == 4 and info[b]["ns_iters"] == 0, (tag, info[b])
    return out, info


# one shape per padded size 16..128: square, tall, and wide (solved transposed);
# at padded sizes 80 and 96 both row-slot counts of A's QR (at most 96 rows
# or more), also for wide inputs
SHAPES = [(5, 5), (16, 16), (17, 9), (9, 17), (32, 32), (40, 33), (33, 100), (48, 48), (128, 47), (64, 64),
          (100, 63), (65, 65), (80, 80), (90, 77), (96, 96), (128, 95), (112, 112), (120, 101), (128, 128),
          (128, 113), (113, 128), (128, 1), (1, 128), (128, 2), (81, 96), (70, 97), (97, 112), (97, 81)]


def test_onesided_uniform_shapes_vs_oracle():
    mats = [uniform_matrix(700 + i, m, n) for i, (m, n) in enumerate(SHAPES)]
    _check_batch(mats, "uniform")


@pytest.mark.parametrize("kexp", [0.0, 4.0, 8.0, 12.0])
def test_onesided_imposed_spectra_vs_oracle(kexp):
    """Log-spaced spectra down to 10^-kexp (the config-2 generator's range)
    and a run of repeated singular values."""
    mats, refs = [], []
    for i, (m, n) in enumerate(SHAPES):
        k = min(m, n)
        s = np.logspace(0.0, -kexp, k) if k > 1 else np.ones(1)
        if k >= 6:
            s[2:5] = s[2]
        s = np.sort(s)[::-1]
        mats.append(imposed_spectrum(7100 + 37 * i + int(kexp), m, n, s))
        refs.append(s)
    _check_batch(mats, f"imposed 1e-{kexp:g}", refs)


def test_onesided_ragged_batch_vs_imposed_and_scaled_plan():
    """512 matrices of the ragged config-2 generator (m ~ U{8..128},
    n ~ U{4..m}, kappa 10^U[0,12], 10% clustered): sigma against the imposed
    spectra and against the scaled reference-iteration plan."""
    mats, refs = [], []
    for i in range(512):
        a, s = workloads.ragged_matrix(workloads.SEED_BATCH + 17, i)
        mats.append(a)
        refs.append(s)
    _say("case ragged 512")
    out, info = svd_batched(mats, precision=OS, handle=_handle(), with_info=True)
    base = svd_batched(mats, precision=LSVD_PREC_F64_SCALED)
    sweeps = []
    for b, (a, (u, s, vt), (_, s2, _)) in enumerate(zip(mats, out, base)):
        n = max(a.shape)
        tag = f"ragged[{b}] {a.shape}"
        assert_factors(a, u, s, vt, 1e-11 * max(1.0, n / 32), tag)
        assert_sigma_close(s, refs[b], n, what=tag + " vs imposed")
        assert_sigma_close(s, s2, n, what=tag + " vs scaled plan")
        assert info[b]["path"] == 4, (tag, info[b])
        sweeps.append(info[b]["jacobi_sweeps"])
    assert max(sweeps) <= 16, max(sweeps)


def test_onesided_degenerate_inputs():
    """Zero, rank-deficient (repeated and zero columns), one column / one
    row, identity and a permuted diagonal with repeated entries."""
    mats = []
    mats.append(np.zeros((20, 12)))
    a = uniform_matrix(801, 40, 30)
    a[:, 7] = a[:, 3]
    a[:, 11] = 0.0
    mats.append(a)
    b = uniform_matrix(802, 64, 10) @ uniform_matrix(803, 10, 64)  # rank 10
    mats.append(b)
    mats.append(uniform_matrix(804, 100, 1))
    mats.append(uniform_matrix(805, 1, 77))
    mats.append(np.eye(33))
    d = np.diag([3.0, 1.0, 3.0, 2.0, 1.0, 0.0, 2.0, 3.0])
    p = rng.uniform(806, 8).argsort()
    mats.append(d[p][:, p[::-1]])
    out, info = _check_batch(mats, "degenerate", resid=1e-11)
    s0 = out[0][1]
    assert np.all(s0 == 0.0)
    assert info[0]["repaired"] == 1 and info[2]["repaired"] in (0, 1)


def test_onesided_nonfinite_input_propagates():
    """A NaN entry: the iteration ends (no rotation passes the test) and
    the singular values come back non-finite, as NaN propagates through
    the reference's arithmetic."""
    a = uniform_matrix(901, 30, 20)
    a[4, 5] = np.nan
    good = uniform_matrix(902, 30, 20)
    out = svd_batched([a, good], precision=OS, handle=_handle())
    assert not np.all(np.isfinite(out[0][1]))
    u, s, vt = out[1]
    assert_factors(good, u, s, vt, 1e-11, "finite neighbour")


def test_onesided_split_routes_large_sizes_to_newton_schulz():
    """With the split at 64 (LSVD_OS_MAX_NP) padded sizes up to 64 take the
    one-sided kernels (path 4) and larger ones the scaled Newton-Schulz
    kernels (path 0); both meet the bar."""
    import os
    from lie_cliffalg_analog_svd_amd import Handle
    old = os.environ.get("LSVD_OS_MAX_NP")
    os.environ["LSVD_OS_MAX_NP"] = "64"
    try:
        h = Handle(0)
    finally:
        if old is None:
            del os.environ["LSVD_OS_MAX_NP"]
        else:
            os.environ["LSVD_OS_MAX_NP"] = old
    mats = [uniform_matrix(950 + i, m, n) for i, (m, n) in enumerate([(70, 60), (100, 100), (40, 33), (128, 90)])]
    out, info = svd_batched(mats, precision=OS, handle=h, with_info=True)
    paths = [d["path"] for d in info]
    assert paths == [4, 0, 4, 0], paths
    for a, (u, s, vt) in zip(mats, out):
        assert_factors(a, u, s, vt, 1e-11 * max(1.0, max(a.shape) / 32), str(a.shape))


def test_onesided_single_solve():
    """lsvd_set_solve_precision(LSVD_PREC_F64_ONESIDED): the single-matrix
    entry points (LieSvdSmall::solve / solve_rectangular) run QR + one-sided
    Jacobi (path 4) and meet the oracle bar; the default handle still runs
    the reference iteration (path 0)."""
    from lie_cliffalg_analog_svd_amd import LSVD_PREC_F64, LieSvdSmall, workloads as wl
    h = _handle_env(LSVD_OS_MAX_NP=128)
    h.set_solve_precision(OS)
    cases = [wl.single_64x64(), uniform_matrix(990, 128, 128), uniform_matrix(991, 100, 37),
             uniform_matrix(992, 20, 90), imposed_spectrum(993, 64, 64, np.logspace(0.0, -12.0, 64))]
    for a in cases:
        _say("single solve", a.shape)
        if a.shape[0] == a.shape[1]:
            u, s, vt = LieSvdSmall.solve(a, handle=h)
        else:
            u, s, vt = LieSvdSmall.solve_rectangular(a, handle=h)
        info = LieSvdSmall.last_info
        assert info["path"] == 4, info
        n = max(a.shape)
        assert_factors(a, u, s, vt, 1e-11 * max(1.0, n / 32), str(a.shape))
        (uo, so, vto), = oracle.svd_batch([a])
        assert_sigma_close(s, so, n, what=f"single {a.shape}")
        if a.shape == (64, 64) and so[-1] > 1e-6 * so[0]:
            vector_parity(u, uo, so, n, "single U")
            vector_parity(vt.T, vto.T, so, n, "single V")
    h.set_solve_precision(LSVD_PREC_F64)
    u, s, vt = LieSvdSmall.solve(cases[0], handle=h)
    assert LieSvdSmall.last_info["path"] == 0, LieSvdSmall.last_info


def test_onesided_power_of_two_scaling_is_exact():
    """||A||_F is formed after bringing the entries to [-1, 1] by the largest
    one's power of two, so A 2^k gives the same U and V bitwise and sigma
    2^k exactly -- also where the squares of the entries overflow (k = 600)
    or underflow (k = -900) in double precision."""
    base = [uniform_matrix(1200, 40, 30), uniform_matrix(1201, 100, 100),
            imposed_spectrum(1202, 64, 48, np.logspace(0.0, -8.0, 48)), uniform_matrix(1203, 17, 70)]
    for k in (600, -900):
        mats = [np.ldexp(a, k) for a in base]
        _say("scaling 2^", k)
        out = svd_batched(base + mats, precision=OS, handle=_handle())
        for i in range(len(base)):
            (u0, s0, vt0), (u1, s1, vt1) = out[i], out[i + len(base)]
            assert np.array_equal(u0, u1) and np.array_equal(vt0, vt1), (k, base[i].shape)
            assert np.array_equal(np.ldexp(s0, k), s1), (k, base[i].shape)
