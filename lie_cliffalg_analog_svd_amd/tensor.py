"""Tensor HO-SVD (src/lie_svd_tensor.rs) on the GPU.

``LieSvdTensor.hosvd3`` factors a 3-way tensor as core x1 U1 x2 U2 x3 U3: the
mode Grams of the tensor, their LieSvdSmall solves (one batched call for
every Gram of every tensor passed together), and the core rotation.
``hosvd3_batched`` takes many tensors of one shape at once.
``tensor_relative_error`` is the reference's host-side diagnostic.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from .handle import Handle, _ptr, default_handle


@dataclass
class TensorPhaseTrace:
    """src/lie_svd_tensor.rs:11-17."""
    shape: tuple
    initial_offdiag: float
    final_offdiag: float
    superdiag_mass_ratio: float


@dataclass
class TensorPhaseSvd3:
    """src/lie_svd_tensor.rs:19-26."""
    core: np.ndarray
    u1: np.ndarray
    u2: np.ndarray
    u3: np.ndarray
    trace: TensorPhaseTrace


def _tensors(x) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    if a.ndim != 4:
        raise _lib.ShapeError(_lib.LSVD_E_SHAPE, "expected (count, n1, n2, n3) tensors")
    return a


def hosvd3_batched(tensors, handle: Handle | None = None) -> list:
    """hosvd3 of every tensor of a (count, n1, n2, n3) array."""
    x = _tensors(tensors)
    count, n1, n2, n3 = x.shape
    if min(n1, n2, n3) == 0:
        raise _lib.ShapeError(_lib.LSVD_E_SHAPE, "empty tensor")
    h = handle or default_handle()
    core = np.zeros_like(x)
    u = [np.zeros((count, n, n)) for n in (n1, n2, n3)]
    tr = (_lib.LsvdTensorTrace * max(count, 1))()
    h.check(_lib.lib().lsvd_hosvd3_f64(h.raw, count, n1, n2, n3, _ptr(x), _ptr(core), _ptr(u[0]), _ptr(u[1]),
                                       _ptr(u[2]), tr))
    return [TensorPhaseSvd3(core[t], u[0][t], u[1][t], u[2][t],
                            TensorPhaseTrace((n1, n2, n3), tr[t].initial_offdiag, tr[t].final_offdiag,
                                             tr[t].superdiag_mass_ratio)) for t in range(count)]


def rotate_tensor3_batched(tensors, q1, q2, q3, handle: Handle | None = None) -> np.ndarray:
    """out[t][a][b][c] = sum q1[t][a][i] q2[t][b][j] q3[t][c][k] x[t][i][j][k]
    (rotate_tensor3, src/lie_svd_tensor.rs:121-145, per tensor)."""
    x = _tensors(tensors)
    count, n1, n2, n3 = x.shape
    qs = [np.ascontiguousarray(np.asarray(q, dtype=np.float64)) for q in (q1, q2, q3)]
    for q, n in zip(qs, (n1, n2, n3)):
        if q.ndim != 3 or q.shape[0] != count or q.shape[2] != n:
            raise _lib.ShapeError(_lib.LSVD_E_SHAPE, "rotate_tensor3: q_m must be (count, r_m, n_m)")
    r1, r2, r3 = (q.shape[1] for q in qs)
    h = handle or default_handle()
    out = np.zeros((count, r1, r2, r3))
    h.check(_lib.lib().lsvd_tensor_rotate3_f64(h.raw, count, n1, n2, n3, r1, r2, r3, _ptr(x), _ptr(qs[0]),
                                               _ptr(qs[1]), _ptr(qs[2]), _ptr(out)))
    return out


class LieSvdTensor:
    @staticmethod
    def hosvd3(tensor, handle: Handle | None = None) -> TensorPhaseSvd3:
        """src/lie_svd_tensor.rs:31-62."""
        t = np.asarray(tensor, dtype=np.float64)
        if t.ndim != 3:
            raise _lib.ShapeError(_lib.LSVD_E_SHAPE, "expected a 3-way tensor")
        return hosvd3_batched(t[None], handle)[0]


def reconstruct_hosvd3(f: TensorPhaseSvd3, handle: Handle | None = None) -> np.ndarray:
    """src/lie_svd_tensor.rs:64-66: rotate_tensor3(core, U1, U2, U3)."""
    return rotate_tensor3_batched(f.core[None], f.u1[None], f.u2[None], f.u3[None], handle)[0]


def tensor_relative_error(a, b) -> float:
    """src/lie_svd_tensor.rs:68-71 (host diagnostic)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.sqrt(np.sum((a - b) ** 2)) / max(np.sqrt(np.sum(a * a)), 1e-300))
