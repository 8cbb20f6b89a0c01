"""Phase-JADE joint diagonalization on the GPU (src/lie_svd_joint.rs).

``LieSvdJoint.diagonalize_symmetric[_with_params]`` mirror the reference:
one family of K symmetric n x n matrices in, ``(V, diagonals, trace)`` out,
with ``work_k = V^T M_k V`` nearly diagonal.  ``diagonalize_families`` runs
many independent families in one call (the batched form of the same
entry point).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from .handle import Handle, _ptr, default_handle


@dataclass
class JointDiagonalizationParams:
    """src/lie_svd_joint.rs:18-35."""
    max_sweeps: int
    tol: float = 1e-12
    min_pair_scale: float = 1e-14
    line_search_steps: int = 6

    @staticmethod
    def for_n(n: int) -> "JointDiagonalizationParams":
        return JointDiagonalizationParams(12 + max(n, 1))

    @staticmethod
    def default() -> "JointDiagonalizationParams":
        return JointDiagonalizationParams.for_n(32)

    def _c(self) -> _lib.LsvdJointParams:
        return _lib.LsvdJointParams(int(self.max_sweeps), float(self.tol), float(self.min_pair_scale),
                                    int(self.line_search_steps))


@dataclass
class JointDiagonalizationTrace:
    """src/lie_svd_joint.rs:43-50."""
    initial_offdiag: float
    final_offdiag: float
    sweeps: int
    rotations: int
    rejected_rotations: int


def diagonalize_families(families, params: JointDiagonalizationParams | None = None,
                         handle: Handle | None = None):
    """families: (F, K, n, n) -> (V (F, n, n), diagonals (F, K, n), traces)."""
    fam = np.ascontiguousarray(np.asarray(families, dtype=np.float64))
    if fam.ndim != 4 or fam.shape[2] != fam.shape[3]:
        raise _lib.ShapeError(_lib.LSVD_E_SHAPE, "LieSvdJoint: matrices must be square")
    F, K, n, _ = fam.shape
    if K == 0:
        raise _lib.ShapeError(_lib.LSVD_E_SHAPE, "LieSvdJoint: empty matrix set")
    params = params or JointDiagonalizationParams.for_n(n)
    h = handle or default_handle()
    v = np.zeros((F, n, n)); d = np.zeros((F, K, n))
    tr = (_lib.LsvdJointTrace * max(F, 1))()
    cp = params._c()
    st = _lib.lib().lsvd_joint_diag_sym_f64(h.raw, F, K, n, _ptr(fam), C.byref(cp), _ptr(v), _ptr(d), tr)
    h.check(st)
    traces = [JointDiagonalizationTrace(**tr[f].as_dict()) for f in range(F)]
    return v, d, traces


@dataclass
class JointSvdTrace:
    """src/lie_svd_joint.rs:53-59."""
    initial_offdiag: float
    final_offdiag: float
    sweeps: int
    rotations: int
    rejected_rotations: int


def svd_families(families, params: JointDiagonalizationParams | None = None, handle: Handle | None = None):
    """Two-sided family SVD of independent families.  families: (F, K, rows,
    cols) -> (U (F, rows, rows), sigmas (F, K, min), Vt (F, cols, cols),
    traces)."""
    fam = np.ascontiguousarray(np.asarray(families, dtype=np.float64))
    if fam.ndim != 4:
        raise _lib.ShapeError(_lib.LSVD_E_SHAPE, "LieSvdJoint: families must be (F, K, rows, cols)")
    F, K, rows, cols = fam.shape
    if K == 0:
        raise _lib.ShapeError(_lib.LSVD_E_SHAPE, "LieSvdJoint: empty matrix set")
    if rows == 0 or cols == 0:
        raise _lib.ShapeError(_lib.LSVD_E_SHAPE, "LieSvdJoint: empty matrix")
    kk = min(rows, cols)
    params = params or JointDiagonalizationParams.for_n(kk)
    h = handle or default_handle()
    u = np.zeros((F, rows, rows)); sg = np.zeros((F, K, kk)); vt = np.zeros((F, cols, cols))
    tr = (_lib.LsvdJointTrace * max(F, 1))()
    cp = params._c()
    st = _lib.lib().lsvd_joint_svd_f64(h.raw, F, K, rows, cols, _ptr(fam), C.byref(cp), _ptr(u), _ptr(sg), _ptr(vt),
                                       tr)
    h.check(st)
    traces = [JointSvdTrace(**tr[f].as_dict()) for f in range(F)]
    return u, sg, vt, traces


class LieSvdJoint:
    @staticmethod
    def diagonalize_symmetric(matrices, handle: Handle | None = None):
        """src/lie_svd_joint.rs:64-69."""
        mats = list(matrices)
        n = mats[0].shape[0] if mats else 0
        return LieSvdJoint.diagonalize_symmetric_with_params(mats, JointDiagonalizationParams.for_n(n), handle)

    @staticmethod
    def diagonalize_symmetric_with_params(matrices, params: JointDiagonalizationParams,
                                          handle: Handle | None = None):
        """src/lie_svd_joint.rs:71-158 -> (V, [diag_k], trace)."""
        mats = [np.asarray(m, dtype=np.float64) for m in matrices]
        if not mats:
            raise _lib.ShapeError(_lib.LSVD_E_SHAPE, "LieSvdJoint: empty matrix set")
        n = mats[0].shape[0]
        for m in mats:
            if m.shape != (n, n):
                raise _lib.ShapeError(_lib.LSVD_E_SHAPE, "LieSvdJoint: size mismatch / not square")
        v, d, tr = diagonalize_families(np.stack(mats)[None], params, handle)
        return v[0], [d[0, k] for k in range(len(mats))], tr[0]

    @staticmethod
    def joint_svd(matrices, handle: Handle | None = None):
        """src/lie_svd_joint.rs:160-168."""
        mats = list(matrices)
        n = min(mats[0].shape) if mats else 0
        return LieSvdJoint.joint_svd_with_params(mats, JointDiagonalizationParams.for_n(n), handle)

    @staticmethod
    def joint_svd_with_params(matrices, params: JointDiagonalizationParams, handle: Handle | None = None):
        """src/lie_svd_joint.rs:170-262 -> (U, [sigmas_k], Vt, trace)."""
        mats = [np.asarray(m, dtype=np.float64) for m in matrices]
        if not mats:
            raise _lib.ShapeError(_lib.LSVD_E_SHAPE, "LieSvdJoint: empty matrix set")
        rows, cols = mats[0].shape
        if rows == 0 or cols == 0:
            raise _lib.ShapeError(_lib.LSVD_E_SHAPE, "LieSvdJoint: empty matrix")
        for m in mats:
            if m.shape[0] != rows:
                raise _lib.ShapeError(_lib.LSVD_E_SHAPE, "LieSvdJoint: row-size mismatch")
            if m.shape[1] != cols:
                raise _lib.ShapeError(_lib.LSVD_E_SHAPE, "LieSvdJoint: col-size mismatch")
        u, sg, vt, tr = svd_families(np.stack(mats)[None], params, handle)
        return u[0], [sg[0, k] for k in range(len(mats))], vt[0], tr[0]
