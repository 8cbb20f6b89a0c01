"""The adaptive router's decision rules on the host (no GPU): the
reference tests' triage structs (src/lie_svd_adaptive.rs:444-511) and
agreement with the oracle's restatement of choose_route on random triages
spread around every threshold."""
from __future__ import annotations

import numpy as np

import oracle
from lie_cliffalg_analog_svd_amd.adaptive import AdaptiveRoute, AdaptiveTriage, LieSvdAdaptiveParams, route_of

FIELDS = ["offdiag_ratio", "diagonal_dominance", "row_cv", "col_cv", "row_col_mismatch", "symmetry_ratio",
          "transpose_torsion_ratio", "diagonal_entropy", "row_mass_entropy", "col_mass_entropy", "phase_mean_stress",
          "phase_max_twist", "phase_causal_disbalance", "phase_entropy_gap", "phase_torsion_energy",
          "phase_chirality_balance", "phase_entropy", "suspicious_score"]


def test_reference_triage_structs():
    chiral = AdaptiveTriage(32, 0.40, 0.30, 0.30, 0.30, 0.10, 0.60, 0.40, 0.70, 0.80, 0.80, 5.0, 0.50, 0.10, 0.05,
                            4.0, 0.45, 0.6, 0.5)
    assert route_of(chiral) == AdaptiveRoute.PhaseFlow
    low = AdaptiveTriage(**{**chiral.__dict__, "phase_chirality_balance": 0.05})
    assert route_of(low) != AdaptiveRoute.PhaseFlow
    degen = AdaptiveTriage(32, 0.98, 0.17, 0.25, 0.25, 0.22, 0.67, 0.74, 0.94, 0.99, 0.99, 10.0, 0.98, 0.1, 0.02,
                           5.0, 0.05, 0.8, 1.58)
    assert oracle.route_predicate("balanced_degenerate", degen.__dict__)
    assert route_of(degen) == AdaptiveRoute.PhaseFlow  # balanced degenerate cases take PhaseFlow first
    assert route_of(AdaptiveTriage(**{**degen.__dict__, "n": 4})) == AdaptiveRoute.Micro
    assert route_of(AdaptiveTriage(**{**degen.__dict__, "n": 600})) == AdaptiveRoute.Hybrid


def test_route_rules_match_oracle():
    rng = np.random.default_rng(12)
    p = LieSvdAdaptiveParams()
    routes = set()
    for _ in range(4000):
        d = {f: float(rng.uniform(0.0, 1.2)) for f in FIELDS}
        d["phase_mean_stress"] = float(rng.choice([0.5, 2.0, 60.0]))
        d["phase_torsion_energy"] = float(rng.choice([0.5, 2.0]))
        d["suspicious_score"] = float(rng.uniform(0.5, 2.0))
        d["offdiag_ratio"] = float(rng.choice([1e-6, rng.uniform(0, 1)]))
        n = int(rng.choice([3, 16, 64, 65, 200, 513]))
        t = AdaptiveTriage(n, *[d[f] for f in FIELDS])
        r_or = oracle.lib().lso_choose_route(oracle.C.byref(oracle._triage(dict(d, n=n))),
                                            oracle.C.byref(oracle.adaptive_params_default()))
        got = route_of(t, p)
        assert int(got) == r_or, (d, n)
        routes.add(int(got))
    assert routes == {0, 1, 2, 3, 4}
