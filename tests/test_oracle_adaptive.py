"""Pins the oracle's adaptive triage and router (src/lie_svd_adaptive.rs,
src/lie_svd_phasehealth.rs, src/lie_svd_block4.rs) on the reference tests'
own matrices and triage structs (adaptive :403-548, phase health :361-463),
plus closed forms: the statistics of diagonal, symmetric and causal-shift
matrices worked by hand, and a numpy restatement-free check of the entropy
and ratio definitions on a 2 x 2 case."""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle


def structured_stress(n):
    """structured_stress_matrix (src/lie_svd_adaptive.rs:341-346)."""
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    return np.where(i // 6 == j // 6, 4.0, 0.02) + np.sin(((i + j) * 31).astype(float)) * 1e-3


def causal_flow(n, up):
    a = np.eye(n)
    a[np.arange(n - 1), np.arange(1, n)] = up
    return a


def lcg_matrix(n, seed=17):
    """The seeded test's generator (:431-442): 64-bit LCG, top 53 bits."""
    state = seed
    out = np.empty(n * n)
    for k in range(n * n):
        state = (state * 6364136223846793005 + 1) % (1 << 64)
        out[k] = 2.0 * ((state >> 11) * (1.0 / (1 << 53))) - 1.0
    return out.reshape(n, n)


def test_reference_route_fixtures():
    n = 16
    assert oracle.choose_route(np.diag(2.0 - np.arange(n) / n))[0] == "Small"           # :403-412
    assert oracle.choose_route(structured_stress(24))[0] == "CoreFlowTopo"              # :414-420
    i, j = np.meshgrid(np.arange(24), np.arange(24), indexing="ij")
    assert oracle.choose_route(np.sin((i * 13 + j * 7 + 3).astype(float)))[0] == "Small"  # :422-428
    assert oracle.choose_route(lcg_matrix(32))[0] == "Small"                            # :430-442
    assert oracle.choose_route(causal_flow(24, 5.0))[0] == "PhaseFlow"                  # :522-534
    assert oracle.choose_route(causal_flow(128, 5.0))[0] == "Small"                     # :536-548
    assert oracle.choose_route(np.eye(4))[0] == "Micro"
    big = np.random.default_rng(3).standard_normal((520, 520))
    assert oracle.choose_route(big)[0] == "Hybrid"                                      # n > small_max_n


def test_reference_triage_predicates():
    base = dict(n=32, offdiag_ratio=0.98, diagonal_dominance=0.17, row_cv=0.25, col_cv=0.25,
                row_col_mismatch=0.22, symmetry_ratio=0.67, transpose_torsion_ratio=0.74, diagonal_entropy=0.94,
                row_mass_entropy=0.99, col_mass_entropy=0.99, phase_mean_stress=10.0, phase_max_twist=0.98,
                phase_causal_disbalance=0.1, phase_entropy_gap=0.02, phase_torsion_energy=5.0,
                phase_chirality_balance=0.05, phase_entropy=0.8, suspicious_score=1.58)
    assert oracle.route_predicate("balanced_degenerate", base)                          # :444-470
    chiral = dict(n=32, offdiag_ratio=0.40, diagonal_dominance=0.30, row_cv=0.30, col_cv=0.30,
                  row_col_mismatch=0.10, symmetry_ratio=0.60, transpose_torsion_ratio=0.40, diagonal_entropy=0.70,
                  row_mass_entropy=0.80, col_mass_entropy=0.80, phase_mean_stress=5.0, phase_max_twist=0.50,
                  phase_causal_disbalance=0.10, phase_entropy_gap=0.05, phase_torsion_energy=4.0,
                  phase_chirality_balance=0.45, phase_entropy=0.6, suspicious_score=0.5)
    assert oracle.route_predicate("phaseflow", chiral)                                  # :472-511
    assert not oracle.route_predicate("phaseflow", dict(chiral, phase_chirality_balance=0.05))


def test_phase_health_fixtures():
    t = oracle.analyze_matrix(causal_flow(12, 3.0))                                     # :387-392
    assert t["phase_causal_disbalance"] > 0.99 and t["phase_max_twist"] > 0.9
    a = causal_flow(12, 3.0)
    direct = float(np.sqrt(np.sum(((a - a.T) * 0.5) ** 2)))                            # :406-421
    assert abs(t["phase_torsion_energy"] - direct) < 1e-12
    i, j = np.meshgrid(np.arange(6), np.arange(6), indexing="ij")
    sym = np.where(i == j, 2.0 + i, 0.1 * (i + j))                                      # :423-438
    assert oracle.analyze_matrix(sym)["phase_torsion_energy"] < 1e-12
    diag = np.diag([4.0, 3.0, 2.0, 1.0, 0.5, 0.25])                                     # :440-455
    dense = 1.0 + 0.3 * np.sin((i * 5 + j * 3).astype(float))
    assert oracle.analyze_matrix(diag)["phase_entropy"] < oracle.analyze_matrix(dense)["phase_entropy"]
    assert 0.0 <= t["phase_chirality_balance"] <= 1.0 and 0.0 <= t["phase_entropy"] <= 1.0


def test_closed_forms():
    """Statistics worked by hand: a diagonal matrix has no off-diagonal mass,
    full diagonal dominance, zero torsion and causal disbalance; a 2 x 2
    [[a, b], [c, d]] has the ratios of its entries."""
    d = np.diag([3.0, 1.0, 2.0, 5.0, 4.0])
    t = oracle.analyze_matrix(d)
    assert t["offdiag_ratio"] == 0.0 and abs(t["diagonal_dominance"] - 1.0) < 1e-15
    assert t["phase_torsion_energy"] == 0.0 and t["phase_causal_disbalance"] == 0.0
    assert abs(t["diagonal_entropy"] - (-sum(p * math.log(p) for p in d.diagonal() / 15) / math.log(5))) < 1e-14
    a = np.array([[1.0, 2.0], [3.0, 4.0]])
    t = oracle.analyze_matrix(a)
    nrm = math.sqrt(30.0)
    assert abs(t["offdiag_ratio"] - math.sqrt(13.0) / nrm) < 1e-15
    assert abs(t["diagonal_dominance"] - math.sqrt(17.0) / nrm) < 1e-15
    # symmetric part 2.5, skew part 0.5 (each counted twice)
    assert abs(t["symmetry_ratio"] - 2.5 / math.sqrt(2.5 ** 2 + 0.25)) < 1e-15
    assert abs(t["phase_causal_disbalance"] - 5.0 / 13.0) < 1e-15
    rl1, cl1 = np.array([3.0, 7.0]), np.array([4.0, 6.0])
    assert abs(t["row_col_mismatch"] - math.sqrt(2.0) / math.sqrt(np.sum(rl1 ** 2 + cl1 ** 2))) < 1e-15
    rl2 = np.array([math.sqrt(5.0), 5.0])
    assert abs(t["row_cv"] - np.std(rl2) / np.mean(rl2)) < 1e-15
