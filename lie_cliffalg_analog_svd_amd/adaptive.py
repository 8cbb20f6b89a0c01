"""Adaptive triage and router (src/lie_svd_adaptive.rs).

``analyze_matrix`` / ``analyze_batch`` compute the reference's matrix
triage -- off-diagonal and diagonal mass, row / column norm spreads, symmetry
and torsion, mass entropies, the row / column phase signature and the Block4
chirality balance -- on the GPU (one workgroup per matrix, a whole batch per
call); ``choose_route`` applies the reference's decision rules.  The
research routes it can name (PhaseFlow, CoreFlowTopo, Hybrid) are not part
of this library: ``LieSvdAdaptive.solve`` reports the reference's route in
its trace and solves those matrices with LieSvdSmall (``solved_by``).
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass, fields

import numpy as np

from . import _lib
from .handle import Handle, _ptr, default_handle
from .solvers import LieSvdMicro, LieSvdSmall


class AdaptiveRoute(enum.IntEnum):
    """src/lie_svd_adaptive.rs:8-15."""
    Micro = 0
    Small = 1
    PhaseFlow = 2
    CoreFlowTopo = 3
    Hybrid = 4


@dataclass
class LieSvdAdaptiveParams:
    """src/lie_svd_adaptive.rs:59-78."""
    small_max_n: int = 512
    coreflow_topo_max_n: int = 64
    phaseflow_max_n: int = 64
    nearly_diagonal_ratio: float = 1e-5
    suspicious_threshold: float = 1.15

    def _c(self) -> _lib.LsvdAdaptiveParams:
        return _lib.LsvdAdaptiveParams(int(self.small_max_n), int(self.coreflow_topo_max_n), int(self.phaseflow_max_n),
                                       0, float(self.nearly_diagonal_ratio), float(self.suspicious_threshold))


@dataclass
class AdaptiveTriage:
    """src/lie_svd_adaptive.rs:18-50."""
    n: int
    offdiag_ratio: float
    diagonal_dominance: float
    row_cv: float
    col_cv: float
    row_col_mismatch: float
    symmetry_ratio: float
    transpose_torsion_ratio: float
    diagonal_entropy: float
    row_mass_entropy: float
    col_mass_entropy: float
    phase_mean_stress: float
    phase_max_twist: float
    phase_causal_disbalance: float
    phase_entropy_gap: float
    phase_torsion_energy: float
    phase_chirality_balance: float
    phase_entropy: float
    suspicious_score: float

    def _c(self) -> _lib.LsvdTriage:
        t = _lib.LsvdTriage()
        for f in fields(self):
            setattr(t, f.name, getattr(self, f.name))
        return t


@dataclass
class AdaptiveTrace:
    """src/lie_svd_adaptive.rs:52-56; solved_by names the solver this
    library used (the route itself for Micro and Small)."""
    route: AdaptiveRoute
    triage: AdaptiveTriage
    solved_by: AdaptiveRoute | None = None


def _triage(t: _lib.LsvdTriage) -> AdaptiveTriage:
    return AdaptiveTriage(int(t.n), *[float(getattr(t, f)) for f in _lib.TRIAGE_FIELDS])


def analyze_batch(mats, params: LieSvdAdaptiveParams | None = None, handle: Handle | None = None) -> list:
    """choose_route (src/lie_svd_adaptive.rs:103-119) for every matrix of a
    batch in one GPU call -> list of AdaptiveTrace."""
    ms = [np.ascontiguousarray(np.asarray(m, dtype=np.float64)) for m in mats]
    for m in ms:
        if m.ndim != 2 or m.shape[0] != m.shape[1]:
            raise _lib.ShapeError(_lib.LSVD_E_SHAPE, "adaptive SVD expects square matrices")
    B = len(ms)
    if B == 0:
        return []
    n = np.array([m.shape[0] for m in ms], dtype=np.int32)
    off = np.zeros(B, dtype=np.int64)
    off[1:] = np.cumsum(n.astype(np.int64) ** 2)[:-1]
    pool = np.concatenate([m.ravel() for m in ms] + [np.zeros(1)])
    out = (_lib.LsvdTriage * B)()
    cp = (params or LieSvdAdaptiveParams())._c()
    h = handle or default_handle()
    h.check(_lib.lib().lsvd_analyze_batched_f64(h.raw, B, n.ctypes.data_as(_lib._i32p), off.ctypes.data_as(_lib._i64p),
                                                _ptr(pool), C.byref(cp), out))
    return [AdaptiveTrace(AdaptiveRoute(out[b].route), _triage(out[b])) for b in range(B)]


def analyze_matrix(mat, handle: Handle | None = None) -> AdaptiveTriage:
    """analyze_matrix (src/lie_svd_adaptive.rs:186-274)."""
    return analyze_batch([mat], handle=handle)[0].triage


def choose_route(mat, params: LieSvdAdaptiveParams | None = None, handle: Handle | None = None) -> AdaptiveTrace:
    """choose_route (src/lie_svd_adaptive.rs:103-119)."""
    return analyze_batch([mat], params, handle)[0]


def route_of(triage: AdaptiveTriage, params: LieSvdAdaptiveParams | None = None) -> AdaptiveRoute:
    """The decision rules alone on a given triage (:103-184)."""
    cp = (params or LieSvdAdaptiveParams())._c()
    return AdaptiveRoute(_lib.lib().lsvd_choose_route(C.byref(triage._c()), C.byref(cp)))


class LieSvdAdaptive:
    @staticmethod
    def solve(mat, handle: Handle | None = None):
        """src/lie_svd_adaptive.rs:83-85."""
        return LieSvdAdaptive.solve_with_trace(mat, LieSvdAdaptiveParams(), handle)[0]

    @staticmethod
    def solve_with_trace(mat, params: LieSvdAdaptiveParams, handle: Handle | None = None):
        """src/lie_svd_adaptive.rs:87-100: Micro and Small as the reference;
        the research routes' matrices are solved by LieSvdSmall."""
        tr = choose_route(mat, params, handle)
        if tr.route == AdaptiveRoute.Micro:
            res = LieSvdMicro.solve(mat, handle)
            tr.solved_by = AdaptiveRoute.Micro
        else:
            res = LieSvdSmall.solve(mat, handle)
            tr.solved_by = AdaptiveRoute.Small
        return res, tr
