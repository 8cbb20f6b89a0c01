"""Host-side rotor-stream export (no GPU): the native phase-schedule-v1
writer against the oracle's independent writer, the reference's JSON tests
(src/lie_svd_compiler.rs:293-377), and the host libm angles of
lsvd_givens_events."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

import oracle
from lie_cliffalg_analog_svd_amd import _lib
from lie_cliffalg_analog_svd_amd.handle import _ptr
from lie_cliffalg_analog_svd_amd.schedule import HardwarePhaseEvent, HardwareSchedule, HardwareTarget, schedule_json


def _real(layer, i, j, theta_l, theta_r, kind):
    """An event as the reference's from_real_phaseflow shapes it
    (src/lie_svd_compiler.rs:59-88: theta = (theta_l + theta_r) / 2, no
    phases, no energies); the producers themselves are out of scope."""
    return HardwarePhaseEvent(layer, i, j, 0.0, 0.0, 0.5 * (theta_l + theta_r), theta_l, theta_r, None, None,
                              "real_phaseflow", kind)


def _sched(events, channels, target):
    return HardwareSchedule("phase-schedule-v1", target, channels, events, [])


def test_real_schedule_json_contains_layers_and_angles():
    s = _sched([_real(2, 0, 1, 0.125, -0.25, "unwrap")], 4, HardwareTarget.MziMesh)
    js = s.to_json_string()
    assert s.total_events() == 1 and s.conflict_free_layer_count() == 3
    assert '"target": "mzi_mesh"' in js and '"kind": "unwrap"' in js


def test_real_schedule_json_names_kinds():
    s = _sched([_real(0, 2, 3, 0.1, -0.1, "causal_antispin")], 4, HardwareTarget.MziMesh)
    assert "causal_antispin" in s.to_json_string()
    s = _sched([_real(1, 0, 2, 0.03, 0.0, "cross_phase_yinyang")], 4, HardwareTarget.MziMesh)
    assert "cross_phase_yinyang" in s.to_json_string()
    s = _sched([_real(0, 0, 1, -0.02, 0.03, "phase_conjugate"), _real(1, 2, 3, 0.04, -0.05, "bottleneck")], 4,
               HardwareTarget.MziMesh)
    js = s.to_json_string()
    assert "phase_conjugate" in js and "bottleneck" in js
    empty = _sched([], 3, HardwareTarget.FpgaRotorMesh)
    assert empty.conflict_free_layer_count() == 0 and empty.total_events() == 0


def _dicts(s: HardwareSchedule):
    return [{"layer": e.layer, "i": e.i, "j": e.j, "phi_l": e.phi_l, "phi_r": e.phi_r, "theta": e.theta,
             "theta_l": e.theta_l, "theta_r": e.theta_r, "energy_before": e.energy_before,
             "energy_after": e.energy_after, "source": e.source, "kind": e.kind} for e in s.events]


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_native_writer_matches_oracle_writer(seed):
    """Random events over every kind, source and special value, with and
    without energies and signs: both writers print the same bytes."""
    rng = np.random.default_rng(seed)
    specials = [0.0, -0.0, 1.0, -1.0, 0.1, 1e-300, 5e-324, 1.7976931348623157e308, 2.0 ** -27, 123456.789,
                float("nan"), float("inf"), -float("inf"), math.pi]
    kinds = ["golden_prespin", "causal_antispin", "cross_phase_yinyang", "phase_conjugate", "bottleneck",
             "phase_jump", "unwrap", "directional", "surgery", "complex_golden_prespin", "complex_diagonal_phase",
             "givens_elimination"]
    sources = ["real_phaseflow", "complex_phaseflow", "orthogonal_givens"]

    def val():
        return specials[rng.integers(len(specials))] if rng.random() < 0.3 else float(rng.standard_normal()
                                                                                        * 10.0 ** rng.integers(-8, 9))
    ev = []
    for k in range(40):
        ev.append(HardwarePhaseEvent(int(rng.integers(0, 1000)), int(rng.integers(0, 64)), int(rng.integers(0, 64)),
                                     val(), val(), val(), val(), val(), val() if rng.random() < 0.4 else None,
                                     val() if rng.random() < 0.4 else None, sources[k % 3], kinds[k % 12]))
    for signs in ([], [1.0, -1.0, float("nan"), 1.0], [-1.0]):
        for target in (HardwareTarget.MziMesh, HardwareTarget.FpgaRotorMesh):
            s = HardwareSchedule("phase-schedule-v1", target, 17, ev, signs)
            assert s.to_json_string() == oracle.schedule_json(target.as_str(), 17, _dicts(s), signs)


def test_complex_schedule_fields():
    # the shape of src/lie_svd_compiler.rs:90-118 (theta_l = theta_r = theta)
    s = _sched([HardwarePhaseEvent(5, 1, 2, 0.25, -0.5, 0.75, 0.75, 0.75, None, None, "complex_phaseflow",
                                   "complex_diagonal_phase")], 3, HardwareTarget.FpgaRotorMesh)
    e = s.events[0]
    assert (e.layer, e.theta_l, e.theta_r, e.source, e.kind) == (5, 0.75, 0.75, "complex_phaseflow",
                                                                 "complex_diagonal_phase")
    assert s.to_json_string() == oracle.schedule_json("fpga_rotor_mesh", 3, _dicts(s), [])


def test_givens_events_use_host_atan2():
    """lsvd_givens_events: layer = index, theta = theta_l = atan2(s, c) from
    the host libm, bit for bit the oracle's (same libm call)."""
    rng = np.random.default_rng(4)
    q, _ = np.linalg.qr(rng.standard_normal((9, 9)))
    g, sg = oracle.givens_schedule(q)
    n = len(g)
    ij = np.ascontiguousarray(g[:, :2].astype(np.int32))
    cs = np.ascontiguousarray(g[:, 2:4])
    arr = (_lib.LsvdPhaseEvent * n)()
    assert _lib.lib().lsvd_givens_events(n, _ptr(ij), _ptr(cs), arr) == _lib.LSVD_OK
    for k in range(n):
        a = arr[k]
        assert (a.layer, a.i, a.j) == (k, int(g[k, 0]), int(g[k, 1]))
        assert a.theta == g[k, 4] and a.theta_l == g[k, 4] and a.theta_r == 0.0 and a.phi_l == 0.0
    s = HardwareSchedule("phase-schedule-v1", HardwareTarget.MziMesh, 9,
                         [HardwarePhaseEvent(a.layer, a.i, a.j, a.phi_l, a.phi_r, a.theta, a.theta_l, a.theta_r,
                                             None, None, "orthogonal_givens", "givens_elimination") for a in arr],
                         list(sg))
    assert s.to_json_string() == oracle.schedule_json("mzi_mesh", 9, oracle.givens_events(g), sg)


def test_writer_rejects_unknown_codes():
    L = _lib.lib()
    assert L.lsvd_schedule_json(7, 1, 0, None, 0, None, None, 0) == C.c_size_t(-1).value
    bad = HardwareSchedule("phase-schedule-v1", HardwareTarget.MziMesh, 1,
                           [HardwarePhaseEvent(0, 0, 1, 0, 0, 0, 0, 0, None, None, "nowhere", "unwrap")], [])
    with pytest.raises(ValueError):
        schedule_json(bad)
    # truncation: cap bytes written including the NUL, full length returned
    s = _sched([_real(0, 0, 1, 0.5, 0.5, "surgery")], 2, HardwareTarget.MziMesh)
    full = s.to_json_string()
    from lie_cliffalg_analog_svd_amd.schedule import _c_events
    buf = C.create_string_buffer(11)
    n = L.lsvd_schedule_json(0, 2, 1, _c_events(s.events), 0, None, buf, 11)
    assert n == len(full) and buf.value.decode() == full[:10]
