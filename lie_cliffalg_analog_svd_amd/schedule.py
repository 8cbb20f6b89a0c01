"""Rotor-stream export: phase-event schedules (src/lie_svd_compiler.rs).

``HardwareSchedule.from_orthogonal_matrix`` decomposes an orthogonal matrix
-- typically the V or U a solver returned -- into the Givens-elimination
events of the reference's "phase-schedule-v1" format; the elimination runs
on the GPU (``lsvd_givens_schedule_f64``) and ``schedules_from_orthogonal``
does a whole batch in one call.  ``to_json_string`` writes the reference's
JSON text byte for byte (``lsvd_schedule_json``), including the
source and kind names the format defines for the PhaseFlow routes' events
(those routes and their event producers are not part of this library).
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .handle import Handle, _ptr, default_handle

FORMAT_VERSION = "phase-schedule-v1"


class HardwareTarget(enum.IntEnum):
    """src/lie_svd_compiler.rs:12-25."""
    MziMesh = _lib.LSVD_TARGET_MZI_MESH
    FpgaRotorMesh = _lib.LSVD_TARGET_FPGA_ROTOR_MESH

    def as_str(self) -> str:
        return "mzi_mesh" if self is HardwareTarget.MziMesh else "fpga_rotor_mesh"


_SOURCES = {"real_phaseflow": 0, "complex_phaseflow": 1, "orthogonal_givens": 2}
_KINDS = ["golden_prespin", "causal_antispin", "cross_phase_yinyang", "phase_conjugate", "bottleneck",
          "phase_jump", "unwrap", "directional", "surgery", "complex_golden_prespin", "complex_diagonal_phase",
          "givens_elimination"]
_KIND_CODE = {k: c for c, k in enumerate(_KINDS)}


@dataclass
class HardwarePhaseEvent:
    """src/lie_svd_compiler.rs:27-41."""
    layer: int
    i: int
    j: int
    phi_l: float
    phi_r: float
    theta: float
    theta_l: float
    theta_r: float
    energy_before: float | None
    energy_after: float | None
    source: str
    kind: str


@dataclass
class HardwareSchedule:
    """src/lie_svd_compiler.rs:43-56."""
    format_version: str
    target: HardwareTarget
    channels: int
    events: list = field(default_factory=list)
    diagonal_signs: list = field(default_factory=list)

    @staticmethod
    def from_orthogonal_matrix(v, target: HardwareTarget, handle: Handle | None = None) -> "HardwareSchedule":
        """src/lie_svd_compiler.rs:144-197: Givens elimination of a square
        matrix (on the GPU); V = G_1^T ... G_m^T diag(diagonal_signs)."""
        v = np.asarray(v, dtype=np.float64)
        if v.ndim != 2 or v.shape[0] != v.shape[1]:
            raise _lib.ShapeError(_lib.LSVD_E_SHAPE, "HardwareSchedule::from_orthogonal_matrix: matrix must be "
                                  f"square, got {v.shape[0]}x{v.shape[1] if v.ndim == 2 else '?'}")
        return schedules_from_orthogonal([v], target, handle=handle)[0]

    def total_events(self) -> int:
        """src/lie_svd_compiler.rs:199-201."""
        return len(self.events)

    def conflict_free_layer_count(self) -> int:
        """src/lie_svd_compiler.rs:203-211: max layer + 1, 0 without events."""
        return max((e.layer for e in self.events), default=-1) + 1

    def to_json_string(self) -> str:
        """src/lie_svd_compiler.rs:213-268 (the native writer)."""
        return schedule_json(self)


def _c_events(events):
    arr = (_lib.LsvdPhaseEvent * max(len(events), 1))()
    for k, e in enumerate(events):
        if e.source not in _SOURCES or e.kind not in _KIND_CODE:
            raise ValueError(f"unknown event source/kind {e.source!r}/{e.kind!r}")
        flags = (1 if e.energy_before is not None else 0) | (2 if e.energy_after is not None else 0)
        arr[k] = _lib.LsvdPhaseEvent(int(e.layer), int(e.i), int(e.j), flags, float(e.phi_l), float(e.phi_r),
                                     float(e.theta), float(e.theta_l), float(e.theta_r),
                                     float(e.energy_before or 0.0), float(e.energy_after or 0.0),
                                     _SOURCES[e.source], _KIND_CODE[e.kind])
    return arr


def schedule_json(s: HardwareSchedule) -> str:
    arr = _c_events(s.events)
    sg = np.ascontiguousarray(np.asarray(s.diagonal_signs, dtype=np.float64).reshape(-1))
    sp = _ptr(sg) if sg.size else None
    L = _lib.lib()
    need = L.lsvd_schedule_json(int(s.target), int(s.channels), len(s.events), arr, int(sg.size), sp, None, 0)
    if need == C.c_size_t(-1).value:
        raise ValueError("lsvd_schedule_json: unknown target, source or kind")
    buf = C.create_string_buffer(need + 1)
    L.lsvd_schedule_json(int(s.target), int(s.channels), len(s.events), arr, int(sg.size), sp, buf, need + 1)
    return buf.raw[:need].decode()


def givens_batch(mats, transpose: bool = False, theta: bool = True, handle: Handle | None = None):
    """Raw batched Givens elimination (host arrays).  mats: a list of square
    matrices (any sizes <= 1024).  Returns (n_events (B,), ij (E, 2),
    cs (E, 2), theta (E,) or None, signs (S,), event offsets (B + 1,), sign
    offsets (B + 1,)); matrix b's events are rows ev_off[b] .. ev_off[b] +
    n_events[b]."""
    ms = [np.asarray(m, dtype=np.float64) for m in mats]
    for m in ms:
        if m.ndim != 2 or m.shape[0] != m.shape[1]:
            raise _lib.ShapeError(_lib.LSVD_E_SHAPE, "HardwareSchedule::from_orthogonal_matrix: matrix must be "
                                  "square")
    B = len(ms)
    d = np.array([m.shape[0] for m in ms], dtype=np.int32)
    sizes = d.astype(np.int64) ** 2
    v_off = np.zeros(B, dtype=np.int64)
    if B:
        v_off[1:] = np.cumsum(sizes)[:-1]
    pool = np.concatenate([m.reshape(-1) for m in ms]) if B else np.zeros(1)
    pool = np.ascontiguousarray(pool if pool.size else np.zeros(1))
    ev_off = np.zeros(B + 1, dtype=np.int64)
    ev_off[1:] = np.cumsum(d.astype(np.int64) * (d.astype(np.int64) - 1) // 2)
    sg_off = np.zeros(B + 1, dtype=np.int64)
    sg_off[1:] = np.cumsum(d.astype(np.int64))
    E, S = int(ev_off[-1]), int(sg_off[-1])
    n_ev = np.zeros(max(B, 1), dtype=np.int32)
    ij = np.zeros((max(E, 1), 2), dtype=np.int32)
    cs = np.zeros((max(E, 1), 2))
    th = np.zeros(max(E, 1)) if theta else None
    sg = np.zeros(max(S, 1))
    h = handle or default_handle()
    st = _lib.lib().lsvd_givens_schedule_f64(
        h.raw, B, d.ctypes.data_as(_lib._i32p), v_off.ctypes.data_as(_lib._i64p), _ptr(pool), int(bool(transpose)),
        _ptr(n_ev), _ptr(ij), _ptr(cs), _ptr(th) if theta else None, _ptr(sg))
    h.check(st)
    return n_ev[:B], ij[:E], cs[:E], (th[:E] if theta else None), sg[:S], ev_off, sg_off


def schedules_from_orthogonal(mats, target: HardwareTarget = HardwareTarget.MziMesh, transpose: bool = False,
                              handle: Handle | None = None) -> list:
    """HardwareSchedule::from_orthogonal_matrix for every matrix of a batch
    (one GPU call).  Event angles are atan2(s, c) of the device's (c, s)
    taken with the host libm, as the reference computes them."""
    n_ev, ij, cs, _, sg, ev_off, sg_off = givens_batch(mats, transpose=transpose, theta=False, handle=handle)
    out = []
    L = _lib.lib()
    for b, m in enumerate(mats):
        d = int(np.asarray(m).shape[0])
        n = int(n_ev[b])
        e0 = int(ev_off[b])
        ij_b = np.ascontiguousarray(ij[e0:e0 + n])
        cs_b = np.ascontiguousarray(cs[e0:e0 + n])
        arr = (_lib.LsvdPhaseEvent * max(n, 1))()
        if n:
            h_ok = L.lsvd_givens_events(n, _ptr(ij_b), _ptr(cs_b), arr)
            if h_ok != _lib.LSVD_OK:
                raise _lib.LsvdError(h_ok, "lsvd_givens_events")
        ev = [HardwarePhaseEvent(int(a.layer), int(a.i), int(a.j), a.phi_l, a.phi_r, a.theta, a.theta_l, a.theta_r,
                                 None, None, "orthogonal_givens", "givens_elimination") for a in arr[:n]]
        signs = [float(x) for x in sg[int(sg_off[b]):int(sg_off[b]) + d]]
        out.append(HardwareSchedule(FORMAT_VERSION, HardwareTarget(target), d, ev, signs))
    return out
