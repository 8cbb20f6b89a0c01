"""GPU parity of the adaptive triage and router (analyze_matrix /
choose_route, src/lie_svd_adaptive.rs:103-274, with the phase signature
and Block4 balance) against the CPU oracle: every triage statistic to 1e-11
(relative, absolute near zero -- the device sums in another order and its
log / atan2 differ in the last bits), the route on the reference tests'
matrices, and LieSvdAdaptive.solve's factors."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from lie_cliffalg_analog_svd_amd import ShapeError
from lie_cliffalg_analog_svd_amd.adaptive import AdaptiveRoute, LieSvdAdaptive, LieSvdAdaptiveParams, analyze_batch
from tests._util import assert_factors
from tests.test_oracle_adaptive import causal_flow, lcg_matrix, structured_stress

pytestmark = pytest.mark.gpu

FIELDS = ["offdiag_ratio", "diagonal_dominance", "row_cv", "col_cv", "row_col_mismatch", "symmetry_ratio",
          "transpose_torsion_ratio", "diagonal_entropy", "row_mass_entropy", "col_mass_entropy", "phase_mean_stress",
          "phase_max_twist", "phase_causal_disbalance", "phase_entropy_gap", "phase_torsion_energy",
          "phase_chirality_balance", "phase_entropy", "suspicious_score"]


def _fixtures():
    i, j = np.meshgrid(np.arange(24), np.arange(24), indexing="ij")
    return [("nearly_diagonal", np.diag(2.0 - np.arange(16) / 16), "Small"),
            ("structured_stress", structured_stress(24), "CoreFlowTopo"),
            ("uniform_sin", np.sin((i * 13 + j * 7 + 3).astype(float)), "Small"),
            ("seeded_lcg", lcg_matrix(32), "Small"),
            ("causal_jordan_24", causal_flow(24, 5.0), "PhaseFlow"),
            ("causal_jordan_128", causal_flow(128, 5.0), "Small"),
            ("micro", np.eye(3), "Micro")]


def test_routes_on_reference_fixtures():
    print("adaptive fixtures", flush=True)
    fx = _fixtures()
    traces = analyze_batch([m for _, m, _ in fx])
    for (name, m, want), tr in zip(fx, traces):
        assert tr.route == AdaptiveRoute[want], (name, tr.route)


def test_triage_statistics_vs_oracle():
    print("adaptive statistics", flush=True)
    rng = np.random.default_rng(600)
    mats = [m for _, m, _ in _fixtures()]
    mats += [rng.standard_normal((n, n)) for n in (1, 2, 5, 8, 33, 64, 100)]
    mats += [np.zeros((6, 6)), np.triu(rng.standard_normal((20, 20))), np.zeros((0, 0))]
    sym = rng.standard_normal((40, 40))
    mats.append(sym + sym.T)
    traces = analyze_batch(mats)
    for k, (m, tr) in enumerate(zip(mats, traces)):
        ref = oracle.analyze_matrix(m) if m.shape[0] > 0 else None
        if ref is None:
            assert tr.route == AdaptiveRoute.Micro and tr.triage.n == 0
            continue
        assert tr.triage.n == m.shape[0]
        for f in FIELDS:
            g, w = getattr(tr.triage, f), ref[f]
            assert abs(g - w) <= 1e-11 * max(abs(w), 1.0), (k, m.shape, f, g, w)
        r, _ = oracle.choose_route(m)
        assert tr.route.name == r, (k, m.shape, tr.route, r)


def test_adaptive_solve_and_errors():
    print("adaptive solve", flush=True)
    i, j = np.meshgrid(np.arange(10), np.arange(10), indexing="ij")
    a = np.sin((i * 13 + j * 7).astype(float))                        # src/lie_svd_adaptive.rs:513-520
    (u, s, vt), tr = LieSvdAdaptive.solve_with_trace(a, LieSvdAdaptiveParams())
    assert_factors(a, u, s, vt, 1e-10, "adaptive")
    assert tr.solved_by in (AdaptiveRoute.Small, AdaptiveRoute.Micro)
    u, s, vt = LieSvdAdaptive.solve(causal_flow(24, 5.0))
    assert_factors(causal_flow(24, 5.0), u, s, vt, 1e-10, "adaptive causal")
    with pytest.raises(ShapeError):
        analyze_batch([np.zeros((3, 4))])
