"""GPU parity of LieSvdSmall / LieSvdMicro / eigh_jacobi_full against the
CPU oracle, through the C ABI.

Tolerances (FP64): singular values two-tier -- relative <= 1e-12 for
sigma_i >= 1e-3 sigma_max and absolute <= 64 n eps sigma_max everywhere;
singular vectors by the largest principal angle per cluster of equal
singular values, bounded by 1e3 n eps sigma_max / gap; residuals (recon,
orth_u, orth_v) at the FP64 level.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle
from lie_cliffalg_analog_svd_amd import (LieSvd, LieSvdMicro, LieSvdSmall, ShapeError, eigh_jacobi_full,
                                         metrics, profiles, rng, svd_batched)
from tests._util import EPS, assert_factors, assert_sigma_close, imposed_spectrum, subspace_sine, uniform_matrix
from tests._util import vector_parity as _vector_parity

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _check_vs_oracle(a, what, rel_tol=1e-12, resid=1e-11):
    _say("case", what, a.shape)
    if a.shape[0] == a.shape[1]:
        u, s, vt = LieSvdSmall.solve(a)
        uo, so, vto, _ = oracle.small_solve(a)
    else:
        u, s, vt = LieSvdSmall.solve_rectangular(a)
        uo, so, vto, _ = oracle.small_solve_rectangular(a)
    n = max(a.shape)
    assert_factors(a, u, s, vt, resid * max(1.0, n / 32), what)
    assert_sigma_close(s, so, n, rel_tol=rel_tol, what=what)
    _vector_parity(u, uo, so, n, what + " U")
    _vector_parity(vt.T, vto.T, so, n, what + " V")
    return u, s, vt


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 9, 13, 16, 17, 31, 32, 33, 48, 63, 64, 65, 96, 127, 128])
def test_square_uniform_vs_oracle(n):
    """src/lie_svd_small.rs:543-569 style inputs (U(-1,1))."""
    _check_vs_oracle(uniform_matrix(100 + n, n, n), f"uniform n={n}")


@pytest.mark.parametrize("n", [129, 150, 200, 256])
def test_square_above_on_chip_sizes_vs_oracle(n):
    """N > 128: the workgroup-per-matrix kernels with global slots
    (LieSvd routes these to LieSvdSmall up to SMALL_MAX_N = 512).  The
    handle keeps single solves off the chip-wide route (LSVD_LARGE_MIN_N
    above every n), which has its own tests."""
    import os
    from lie_cliffalg_analog_svd_amd import Handle
    old = os.environ.get("LSVD_LARGE_MIN_N")
    os.environ["LSVD_LARGE_MIN_N"] = "100000"
    try:
        h = Handle()
    finally:
        if old is None:
            del os.environ["LSVD_LARGE_MIN_N"]
        else:
            os.environ["LSVD_LARGE_MIN_N"] = old
    a = uniform_matrix(300 + n, n, n)
    _say("case wg", n)
    u, s, vt = LieSvdSmall.solve(a, h)
    assert LieSvdSmall.last_info["path"] == 0
    uo, so, vto, _ = oracle.small_solve(a)
    assert_factors(a, u, s, vt, 1e-11 * max(1.0, n / 32), f"wg n={n}")
    assert_sigma_close(s, so, n, what=f"wg n={n}")
    _vector_parity(u, uo, so, n, f"wg n={n} U")
    _vector_parity(vt.T, vto.T, so, n, f"wg n={n} V")
    h.close()


def test_single_64x64_normal():
    """The 64 x 64 N(0,1) single-matrix workload."""
    a = rng.normal_matrix(0x5EED0001, 64, 64)
    _check_vs_oracle(a, "normal64")
    assert LieSvdSmall.last_info["polar_ok"] == 1
    assert 10 <= LieSvdSmall.last_info["ns_iters"] <= 40


@pytest.mark.parametrize("profile", list(profiles.Profile))
@pytest.mark.parametrize("n", [32, 64])
def test_profiles_vs_oracle(profile, n):
    case = profiles.generate(n, profile, 17)
    a = case.a
    _say("profile", profile.value, n)
    u, s, vt = LieSvdSmall.solve(a)
    uo, so, vto, _ = oracle.small_solve(a)
    m = metrics.compute(a, u, s, vt)
    assert m.orth_u < 1e-11 and m.orth_v < 1e-11 and m.rel_recon < 1e-11, m
    assert_sigma_close(s, so, n, rel_tol=1e-12, what=profile.value)
    if profile != profiles.Profile.ExtremeIllConditioned:
        _vector_parity(vt.T, vto.T, so, n, profile.value + " V")


@pytest.mark.parametrize("n", [16, 64])
def test_pei_closed_form_gpu(n):
    a = profiles.pei_matrix(n, 0.01)
    u, s, vt = LieSvdSmall.solve(a)
    assert abs(s[0] - (n + 0.01)) / (n + 0.01) < 1e-13
    assert np.max(np.abs(s[1:] - 0.01)) < 8 * n * EPS * (n + 0.01)
    m = metrics.compute(a, u, s, vt, profiles.pei_matrix_singular_values(n, 0.01))
    assert m.orth_u < 1e-12 and m.orth_v < 1e-12 and m.rel_recon < 1e-13


def test_ill_conditioned_1e15():
    """src/lie_svd_small.rs:575-616 (kappa 1e15, residuals < 1e-6)."""
    n = 32
    sig = np.array([10.0 ** (-15.0 * i / (n - 1)) for i in range(n)])
    a = imposed_spectrum(42, n, n, sig)
    u, s, vt = LieSvdSmall.solve(a)
    m = metrics.compute(a, u, s, vt)
    assert m.rel_recon < 1e-6 and m.orth_u < 1e-6 and m.orth_v < 1e-6
    assert np.max(np.abs(s - sig)) < 1e-13


@pytest.mark.parametrize("m,n", [(30, 3), (20, 15), (4, 11), (64, 60), (100, 95), (128, 4), (9, 128),
                                 (128, 127), (77, 64), (128, 100), (200, 150)])
def test_rectangular_vs_oracle(m, n):
    _check_vs_oracle(uniform_matrix(2026 + m * n, m, n), f"rect {m}x{n}")


def test_rank_deficient_duplicate_column():
    """src/lie_svd_small.rs:517-540.  The GPU front end is Householder QR,
    so U stays orthonormal where the reference's MGS leaves a zero Q
    column; sigma and the reconstruction agree."""
    k = 20
    i = np.arange(k)
    a = np.zeros((k, 3))
    a[:, 0] = np.sin((i * 11 + 3) * 0.23) * 2.0
    a[:, 1] = np.cos((i * 13 + 7) * 0.41)
    a[:, 2] = a[:, 0]
    u, s, vt = LieSvdSmall.solve_rectangular(a)
    assert np.all(np.isfinite(u)) and np.all(np.isfinite(s)) and np.all(np.isfinite(vt))
    m = metrics.compute(a, u, s, vt)
    assert m.rel_recon < 1e-13 and m.orth_u < 1e-13
    _, so, _, _ = oracle.small_solve_rectangular(a)
    assert np.max(np.abs(s - so)) < 64 * 20 * EPS * so[0]


def test_zero_and_empty():
    u, s, vt = LieSvdSmall.solve(np.zeros((0, 0)))
    assert u.shape == (0, 0) and s.shape == (0,)
    u, s, vt = LieSvdSmall.solve(np.zeros((5, 5)))
    assert np.all(s == 0) and np.all(np.isfinite(u)) and np.all(np.isfinite(vt))
    with pytest.raises(ShapeError):
        LieSvdSmall.solve_rectangular(np.zeros((0, 4)))


def test_nonfinite_input_does_not_crash():
    """NaN propagates (the reference has no error path); the GPU returns."""
    a = uniform_matrix(3, 6, 6)
    a[1, 2] = np.nan
    u, s, vt = LieSvdSmall.solve(a)
    assert u.shape == (6, 6)
    assert LieSvdSmall.last_info["polar_ok"] == 0


@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_micro_vs_oracle(n):
    """src/lie_svd_micro.rs:239-251; the same fixed schedule, so factors
    agree entry-wise up to the last bits of atan2/sincos."""
    for b in range(32):
        a = uniform_matrix(900 + n * 1000 + b, n, n, -2.0, 2.0)
        u, s, vt = LieSvdMicro.solve(a)
        uo, so, vto, io = oracle.micro_solve(a)
        assert_factors(a, u, s, vt, 1e-13, f"micro n={n} b={b}")
        assert np.max(np.abs(s - so)) <= 16 * EPS * max(so[0], 1e-300)
        if io["micro_fallback"] == 0 and (n == 1 or np.min(np.abs(np.diff(so))) > 1e-6 * so[0]):
            assert np.max(np.abs(u - uo)) < 1e-12 and np.max(np.abs(vt - vto)) < 1e-12


def test_micro_degenerate_and_router():
    a = np.zeros((4, 4))
    a[0, 0] = a[1, 1] = 10.0
    a[2, 2] = a[3, 3] = 1e-12
    a[0, 1] = 0.25
    a[2, 3] = -0.5
    u, s, vt = LieSvdMicro.solve(a)
    assert_factors(a, u, s, vt, 1e-10)
    u2, s2, vt2 = LieSvd.solve(a)  # n <= 4 routes to Micro (src/lie_svd_adaptive.rs:93)
    assert np.array_equal(s, s2)
    u3, s3, vt3 = LieSvd.solve(uniform_matrix(8, 9, 9))
    assert u3.shape == (9, 9)


def test_micro_escalation_to_small(monkeypatch):
    """LieSvdMicro's fallback to LieSvdSmall::solve (src/lie_svd_micro.rs:
    90-92).  No finite input found reaches it within the reference's 24
    sweeps (the cyclic two-sided rotations converge quadratically), so the
    sweep limit of n = 3, 4 is lowered from the host (LSVD_MICRO_MAX_SWEEPS,
    read when a handle is created; the oracle's sweep_cap is the same hook):
    after one sweep most 3 x 3 and 4 x 4 matrices are not converged and go
    through the batch plan's fallback queue, drained by a LieSvdSmall launch.
    The escalated set, the info path (2) and sigma must match the oracle's,
    and 2 x 2 (one rotor, never escalates) stays on the micro path."""
    from lie_cliffalg_analog_svd_amd import Handle, svd_batched
    monkeypatch.setenv("LSVD_MICRO_MAX_SWEEPS", "1")
    h = Handle(0)
    mats = [uniform_matrix(4400 + b, 2 + b % 3, 2 + b % 3, -2.0, 2.0) for b in range(48)]
    out, info = svd_batched(mats, handle=h, with_info=True)
    escalated = 0
    for b, a in enumerate(mats):
        uo, so, vto, io = oracle.micro_solve(a, sweep_cap=1)
        u, s, vt = out[b]
        what = f"micro capped b={b} n={a.shape[0]}"
        assert info[b]["micro_fallback"] == io["micro_fallback"], what
        assert info[b]["path"] == (2 if io["micro_fallback"] else 1), what
        assert_factors(a, u, s, vt, 1e-12, what)
        assert_sigma_close(s, so, a.shape[0], what=what)
        escalated += io["micro_fallback"]
        if io["micro_fallback"]:
            assert info[b]["ns_iters"] > 0, what  # the LieSvdSmall path ran
    h.close()
    assert escalated >= 16  # the fallback queue really carried work
    # the same matrices with the reference's limit: nothing escalates
    out2, info2 = svd_batched(mats, with_info=True)
    assert all(i["path"] == 1 for i in info2)


def test_micro_sign_convention():
    u, s, vt = LieSvdMicro.solve(np.diag([1.0, -5.0, 2.0]))
    assert np.array_equal(s, [5.0, 2.0, 1.0])
    assert np.array_equal(u, [[0, 0, 1], [-1, 0, 0], [0, 1, 0]])
    assert np.array_equal(vt, [[0, 1, 0], [0, 0, 1], [1, 0, 0]])


@pytest.mark.parametrize("t,u", [(1.0, 4.0), (1.0, 0.0), (2.3, -3.1), (0.5, 10.0)])
def test_eigh_hubbard(t, u):
    h = profiles.hubbard_dimer_hamiltonian(t, u)
    v, e = eigh_jacobi_full(h)
    want = np.sort(profiles.hubbard_dimer_eigenvalues(t, u))[::-1]
    assert np.max(np.abs(e - want)) < 1e-13 * max(1.0, abs(u), abs(t))
    assert np.linalg.norm(v.T @ h @ v - np.diag(e)) < 1e-13 * np.linalg.norm(h)


@pytest.mark.parametrize("n", [5, 30, 64, 100])
def test_eigh_vs_oracle(n):
    g = rng.normal_matrix(21 + n, n, n)
    p = g + g.T
    v, e = eigh_jacobi_full(p)
    vo, eo, _ = oracle.eigh_jacobi_full(p)
    assert np.max(np.abs(e - eo)) < 64 * n * EPS * np.abs(eo).max()
    assert np.linalg.norm(v.T @ v - np.eye(n)) < 1e-12
    assert np.linalg.norm(v.T @ p @ v - np.diag(e)) < 1e-12 * np.linalg.norm(p)


def test_batched_ragged_vs_oracle():
    """A few hundred ragged matrices in one plan (micro + small buckets)."""
    from lie_cliffalg_analog_svd_amd import workloads
    mats, refs = [], []
    for i in range(160):
        a, s = workloads.ragged_matrix(77, i, kmax=6.0)
        mats.append(a)
        refs.append(s)
    for i in range(96):
        mats.append(workloads.tiny_matrix(77, i))
        refs.append(None)
    mats.append(uniform_matrix(5, 7, 19))  # a wide one
    refs.append(None)
    _say("batch", len(mats))
    out, info = svd_batched(mats, with_info=True)
    ora = oracle.svd_batch(mats)
    for b, (a, (u, s, vt), (uo, so, vto)) in enumerate(zip(mats, out, ora)):
        n = max(a.shape)
        assert_factors(a, u, s, vt, 1e-11 * max(1.0, n / 32), f"batch {b} {a.shape}")
        assert_sigma_close(s, so, n, rel_tol=1e-12, what=f"batch {b} {a.shape}")
        if refs[b] is not None:
            assert_sigma_close(s, refs[b], n, rel_tol=1e-12, what=f"imposed {b}")
        assert info[b]["path"] in (0, 1, 2)
        if a.shape[0] == a.shape[1] and a.shape[0] <= 4:
            assert info[b]["path"] in (1, 2)


def test_batched_plan_replay_is_deterministic():
    """A captured plan replayed several times gives bit-identical output."""
    from lie_cliffalg_analog_svd_amd import BatchPlan, default_handle, workloads
    mats = [workloads.ragged_matrix(5, i, kmax=3.0)[0] for i in range(40)]
    mats += [workloads.tiny_matrix(5, i) for i in range(40)]
    plan = BatchPlan(default_handle(), [a.shape[0] for a in mats], [a.shape[1] for a in mats])
    L = plan.layout
    a_pool = workloads.pack(mats)
    outs = []
    for rep in range(3):
        u = np.zeros(L["u_len"]); s = np.zeros(L["s_len"]); vt = np.zeros(L["vt_len"])
        plan.execute(a_pool, u, s, vt)
        outs.append((u, s, vt))
    for rep in (1, 2):
        for x, y in zip(outs[0], outs[rep]):
            assert np.array_equal(x, y)
    plan.close()


def test_batched_scaled_schedule_vs_oracle():
    """LSVD_PREC_F64_SCALED: the same outputs as the reference iteration (to
    the FP64 parity bar) on ragged, ill-conditioned (kappa up to 1e12),
    clustered, rank-deficient and gapped spectra, in fewer NS steps."""
    from lie_cliffalg_analog_svd_amd import LSVD_PREC_F64_SCALED, workloads
    mats, refs = [], []
    for i in range(120):
        a, s = workloads.ragged_matrix(91, i, kmax=12.0)
        mats.append(a)
        refs.append(s)
    r = np.random.default_rng(3)
    for n in (16, 48, 100):  # a gap: three tiny singular values far below the rest
        s = np.concatenate([np.logspace(0, -1, n - 3), [1e-8, 1e-9, 1e-10]])
        mats.append(imposed_spectrum(700 + n, n + 7, n, s))
        refs.append(s)
    dup = r.standard_normal((40, 12))
    dup[:, 5] = dup[:, 2]  # rank deficient
    mats.append(dup)
    refs.append(None)
    _say("batch scaled", len(mats))
    out, info = svd_batched(mats, precision=LSVD_PREC_F64_SCALED, with_info=True)
    plain, info0 = svd_batched(mats, with_info=True)
    ora = oracle.svd_batch(mats)
    for b, (a, (u, s, vt), (uo, so, vto)) in enumerate(zip(mats, out, ora)):
        n = max(a.shape)
        what = f"scaled {b} {a.shape}"
        assert_factors(a, u, s, vt, 1e-11 * max(1.0, n / 32), what)
        if refs[b] is not None:
            # the imposed spectrum is the truth; the oracle's own MGS front end
            # (qr_reduce) loses ~eps kappa sigma_max on tall ill-conditioned
            # inputs, so against the oracle the bound is widened by its error
            truth = np.sort(refs[b])[::-1]
            assert_sigma_close(s, truth, n, rel_tol=1e-12, what="imposed " + what)
            o_err = float(np.max(np.abs(so - truth)) / truth[0])
            g_err = float(np.max(np.abs(s - truth)) / truth[0])
            assert g_err <= max(64 * n * EPS, 2.0 * o_err), what
        else:
            assert_sigma_close(s, so, n, rel_tol=1e-12, what=what)
            _vector_parity(vt.T, vto.T, so, n, what + " V")
    its = np.array([d["ns_iters"] for d in info])
    its0 = np.array([d["ns_iters"] for d in info0])
    _say("ns steps scaled", its.sum(), "plain", its0.sum())
    assert its.sum() < 0.85 * its0.sum()


def test_batch_rejects_low_precision():
    """Batch plans compute in FP64; a low-precision request is an error."""
    from lie_cliffalg_analog_svd_amd import LSVD_PREC_BF16, LSVD_PREC_F32, LsvdError, workloads
    mats = [workloads.ragged_matrix(5, i)[0] for i in range(3)]
    for prec in (LSVD_PREC_F32, LSVD_PREC_BF16):
        with pytest.raises(LsvdError):
            svd_batched(mats, precision=prec)


def test_autotune_plan_knobs_and_results():
    """lsvd_autotune_batch on a ragged batch (512 ragged + 256 tiny): the
    report names valid knobs, times at least the default plan and one
    candidate per knob, and the tuned plan is no slower than the default by
    its own measurement; the tuned plan's sigma equal the default plan's
    (the knobs change scheduling and threads per workgroup only) to the
    two-tier FP64 bar."""
    from lie_cliffalg_analog_svd_amd import LSVD_PREC_F64_SCALED, workloads
    from lie_cliffalg_analog_svd_amd.handle import Handle
    batch = workloads.ragged_batch(512, 256)
    mats = batch.mats
    ms = np.array([a.shape[0] for a in mats], dtype=np.int32)
    ns = np.array([a.shape[1] for a in mats], dtype=np.int32)
    h0 = Handle()
    h1 = Handle()
    try:
        r0 = svd_batched(mats, precision=LSVD_PREC_F64_SCALED, handle=h0)
        rep = h1.autotune(ms, ns, LSVD_PREC_F64_SCALED)
        _say("autotune report", rep)
        assert rep["multi_min_np"] in (64, 80, 96, 129)
        assert 1 <= rep["max_lanes"] <= 4 and 0 <= rep["wide_mask"] < 16
        assert rep["evaluations"] >= 8
        assert 0.0 < rep["tuned_ms"] <= rep["default_ms"]
        r1 = svd_batched(mats, precision=LSVD_PREC_F64_SCALED, handle=h1)
        for b in range(len(mats)):
            k = min(mats[b].shape)
            assert_sigma_close(r1[b][1], r0[b][1], k, what=f"tuned vs default {b}")
    finally:
        h0.close()
        h1.close()
