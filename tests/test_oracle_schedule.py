"""Pins the oracle's HardwareSchedule restatement (src/lie_svd_compiler.rs):
the Givens export (:144-197) and the JSON writer (:213-268).

The elimination is pinned by what the mathematics fixes: exact
reconstruction V = G_1^T ... G_m^T diag(signs) (the reference's own round
trip test, :379-433), the closed form of the first column's angles, the
signs of a generic input ([1, ..., 1, det V]), the skip rule on exact zeros,
and hand-worked 2 x 2 cases.  The JSON text is pinned by hand-written
expected strings (Rust's {:.17e} / {} rules), the reference tests' own
assertions (:293-377), and Python's independent correctly rounded "%.17e".
"""
from __future__ import annotations

import json
import math
import re

import numpy as np
import pytest

import oracle


def haar(rng, d):
    q, r = np.linalg.qr(rng.standard_normal((d, d)))
    return q * np.sign(np.diag(r))[None, :]


def reconstruct(ev, signs):
    """The reference test's reconstruction (:401-415): start from
    diag(signs), apply each event's inverse rotation in reverse order."""
    d = len(signs)
    r = np.diag(np.asarray(signs, dtype=np.float64))
    for i, j, c, s, th in ev[::-1]:
        i, j = int(i), int(j)
        ct, st = math.cos(th), math.sin(th)
        rj, ri = r[j].copy(), r[i].copy()
        r[j] = ct * rj - st * ri
        r[i] = st * rj + ct * ri
    return r


@pytest.mark.parametrize("d", [1, 2, 3, 5, 8, 17, 40])
def test_givens_round_trip_and_order(d):
    rng = np.random.default_rng(211 + d)
    v = haar(rng, d)
    ev, sg = oracle.givens_schedule(v)
    assert len(ev) == d * (d - 1) // 2 and len(sg) == d
    assert [(int(a), int(b)) for a, b in ev[:, :2]] == [(i, j) for j in range(d) for i in range(j + 1, d)]
    assert np.max(np.abs(reconstruct(ev, sg) - v)) < 1e-13 * max(d, 1)
    # c, s are the cosine and sine of theta, and a unit pair
    assert np.allclose(ev[:, 2], np.cos(ev[:, 4]), atol=4e-16) and np.allclose(ev[:, 3], np.sin(ev[:, 4]), atol=4e-16)
    assert np.allclose(ev[:, 2] ** 2 + ev[:, 3] ** 2, 1.0, atol=4e-16)


def test_givens_first_column_closed_form():
    """Column 0: the pivot after eliminating rows 1..i-1 is
    +sqrt(v_00^2 + ... + v_(i-1)0^2), so theta_(i,0) = atan2(v_i0, that)
    for i >= 2 and theta_(1,0) = atan2(v_10, v_00)."""
    rng = np.random.default_rng(5)
    for d in (3, 9, 30):
        v = haar(rng, d)
        ev, _ = oracle.givens_schedule(v)
        col = v[:, 0]
        assert abs(ev[0, 4] - math.atan2(col[1], col[0])) < 4e-16 * math.pi
        for k in range(2, d):
            want = math.atan2(col[k], math.sqrt(float(np.sum(col[:k] ** 2))))
            assert abs(ev[k - 1, 4] - want) < 1e-14, (d, k)


def test_givens_signs_generic_input():
    """After a column's rotations its pivot is a positive norm, so only the
    last column keeps a sign: signs = [1, ..., 1, det V]."""
    rng = np.random.default_rng(6)
    for d in (2, 4, 7, 16):
        for flip in (False, True):
            v = haar(rng, d)
            if flip:
                v[:, 0] *= -1
            _, sg = oracle.givens_schedule(v)
            assert np.array_equal(sg[:-1], np.ones(d - 1))
            assert sg[-1] == np.sign(np.linalg.det(v))


def test_givens_two_by_two_cases():
    t = 0.7
    rot = np.array([[math.cos(t), -math.sin(t)], [math.sin(t), math.cos(t)]])
    ev, sg = oracle.givens_schedule(rot)
    assert len(ev) == 1 and abs(ev[0, 4] - t) < 1e-15 and np.array_equal(sg, [1.0, 1.0])
    refl = np.array([[math.cos(t), math.sin(t)], [math.sin(t), -math.cos(t)]])
    ev, sg = oracle.givens_schedule(refl)
    assert len(ev) == 1 and abs(ev[0, 4] - t) < 1e-15 and np.array_equal(sg, [1.0, -1.0])
    swap = np.array([[0.0, 1.0], [1.0, 0.0]])  # a = 0, b = 1: c = 0, s = 1 exactly
    ev, sg = oracle.givens_schedule(swap)
    assert ev[0, 2] == 0.0 and ev[0, 3] == 1.0 and ev[0, 4] == math.pi / 2 and np.array_equal(sg, [1.0, -1.0])


def test_givens_skip_rule_and_diagonal_inputs():
    rng = np.random.default_rng(7)
    for diag in ([1.0, 1.0, 1.0], [1.0, -1.0, -1.0, 1.0], [-1.0]):
        ev, sg = oracle.givens_schedule(np.diag(diag))
        assert len(ev) == 0 and np.array_equal(sg, diag)
    q1, q2 = haar(rng, 3), haar(rng, 4)
    v = np.zeros((7, 7))
    v[:3, :3] = q1
    v[3:, 3:] = q2
    ev, sg = oracle.givens_schedule(v)
    assert len(ev) == 3 + 6  # exact zeros between the blocks are skipped (:158-160)
    assert all((i < 3) == (j < 3) for i, j in ev[:, :2])
    assert np.max(np.abs(reconstruct(ev, sg) - v)) < 1e-14
    # the threshold itself: |b| <= 1e-300 is skipped, above it is not
    for b, want in ((1e-300, 0), (-1e-300, 0), (2e-300, 1)):
        ev, _ = oracle.givens_schedule(np.array([[1.0, 0.0], [b, 1.0]]))
        assert len(ev) == want, b


def test_givens_nan_propagates_to_signs():
    v = np.eye(3)
    v[2, 0] = np.nan
    ev, sg = oracle.givens_schedule(v)
    assert np.isnan(sg[0]) and np.all(np.isnan(ev[1:, 2:]))


def rust_exp17(x: float) -> str:
    """Python's own correctly rounded %.17e with Rust's exponent style."""
    m, e = ("%.17e" % x).split("e")
    return f"{m}e{int(e)}"


def test_json_hand_written_case():
    ev = [{"layer": 2, "i": 0, "j": 1, "phi_l": 0.0, "phi_r": 0.0, "theta": -0.0625, "theta_l": 0.125,
           "theta_r": -0.25, "source": "real_phaseflow", "kind": "unwrap"},
          {"layer": 3, "i": 1, "j": 2, "phi_l": 1.0, "phi_r": -2.5, "theta": 1e-300, "theta_l": 1e300,
           "theta_r": 5e-324, "energy_before": 3.0, "energy_after": 0.5, "source": "complex_phaseflow",
           "kind": "complex_diagonal_phase"}]
    got = oracle.schedule_json("fpga_rotor_mesh", 4, ev, [1.0, -1.0])
    want = (
        '{\n'
        '  "format_version": "phase-schedule-v1",\n'
        '  "target": "fpga_rotor_mesh",\n'
        '  "channels": 4,\n'
        '  "total_events": 2,\n'
        '  "events": [\n'
        '    {"layer": 2, "i": 0, "j": 1, "phi_l": 0.00000000000000000e0, "phi_r": 0.00000000000000000e0, '
        '"theta": -6.25000000000000000e-2, "theta_l": 1.25000000000000000e-1, "theta_r": -2.50000000000000000e-1, '
        '"source": "real_phaseflow", "kind": "unwrap"},\n'
        '    {"layer": 3, "i": 1, "j": 2, "phi_l": 1.00000000000000000e0, "phi_r": -2.50000000000000000e0, '
        '"theta": 1.00000000000000003e-300, "theta_l": 1.00000000000000005e300, '
        '"theta_r": 4.94065645841246544e-324, "source": "complex_phaseflow", "kind": "complex_diagonal_phase", '
        '"energy_before": 3.00000000000000000e0, "energy_after": 5.00000000000000000e-1}\n'
        '  ],\n'
        '  "diagonal_signs": [1, -1]\n'
        '}\n')
    assert got == want
    # no signs: the key is omitted; no events: an empty list on two lines
    assert oracle.schedule_json("mzi_mesh", 0, [], []) == (
        '{\n  "format_version": "phase-schedule-v1",\n  "target": "mzi_mesh",\n  "channels": 0,\n'
        '  "total_events": 0,\n  "events": [\n  ]\n}\n')


def test_json_reference_assertions_and_round_trip():
    """The reference tests' assertions on real PhaseFlow schedules
    (:293-377), and every number of a Givens schedule parses back to the
    same double and matches Python's independent %.17e."""
    ev = [{"layer": 2, "i": 0, "j": 1, "phi_l": 0.0, "phi_r": 0.0, "theta": 0.5 * (0.125 - 0.25), "theta_l": 0.125,
           "theta_r": -0.25, "source": "real_phaseflow", "kind": "unwrap"}]
    js = oracle.schedule_json("mzi_mesh", 4, ev, [])
    assert '"target": "mzi_mesh"' in js and '"kind": "unwrap"' in js and '"total_events": 1' in js
    assert "diagonal_signs" not in js
    rng = np.random.default_rng(8)
    v = haar(rng, 6)
    g, sg = oracle.givens_schedule(v)
    js = oracle.schedule_json("mzi_mesh", 6, oracle.givens_events(g), sg)
    doc = json.loads(js)
    assert doc["total_events"] == len(g) == 15 and doc["channels"] == 6
    for k, e in enumerate(doc["events"]):
        assert e["layer"] == k and e["theta"] == g[k, 4] and e["theta_l"] == g[k, 4] and e["theta_r"] == 0.0
        assert e["source"] == "orthogonal_givens" and e["kind"] == "givens_elimination"
    assert doc["diagonal_signs"] == list(sg)
    nums = re.findall(r'"theta": ([^,]+),', js)
    assert nums == [rust_exp17(t) for t in g[:, 4]]
