"""Reference-shaped solver API over the gfx950 library.

Names, argument meaning, output conventions and error behaviour follow the
reference crate (skovnats/lie_cliffalg_analog_svd v0.44.0):

- ``LieSvdSmall.solve``             src/lie_svd_small.rs:296-323
- ``LieSvdSmall.solve_rectangular`` src/lie_svd_small.rs:340-360
- ``eigh_jacobi_full``              src/lie_svd_small.rs:212-285
- ``LieSvdMicro.solve``             src/lie_svd_micro.rs:21-52
- ``LieSvd.solve``                  src/lie_svd.rs:49-51 (Micro/Small routing)
- ``run_small``                     src/lib.rs:56-60
- ``digital_polish[_batched]``      src/lie_svd_phaseflow.rs:643-658 (finishing step)

Every function returns ``(U, sigma, Vt)`` as numpy float64 arrays with
sigma descending; shape misuse raises ``ShapeError`` (an AssertionError),
where the reference panics.  All arithmetic runs in HIP kernels.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib
from .handle import BatchPlan, Handle, _ptr, default_handle, info_array, info_dicts, pack_offsets

MICRO_MAX_N = 4      # src/lie_svd.rs:31
SMALL_MAX_N = 512    # src/lie_svd.rs:38


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _call(h: Handle, fn, *args):
    info = _lib.LsvdInfo()
    st = fn(h.raw, *args, C.byref(info))
    h.check(st)
    return info.as_dict()


class LieSvdSmall:
    """FP64 polar (Newton-Schulz) + Jacobi SVD, on the GPU."""

    last_info: dict = {}

    @staticmethod
    def solve(mat, handle: Handle | None = None):
        a = _f64(mat)
        if a.ndim != 2 or a.shape[0] != a.shape[1]:
            raise _lib.ShapeError(_lib.LSVD_E_SHAPE, "LieSvdSmall::solve: matrix must be square N×N")
        n = a.shape[0]
        h = handle or default_handle()
        u = np.zeros((n, n)); s = np.zeros(n); vt = np.zeros((n, n))
        LieSvdSmall.last_info = _call(h, _lib.lib().lsvd_small_solve_f64, _ptr(a), n, _ptr(u), _ptr(s), _ptr(vt))
        return u, s, vt

    @staticmethod
    def solve_rectangular(mat, handle: Handle | None = None):
        a = _f64(mat)
        if a.ndim != 2:
            raise _lib.ShapeError(_lib.LSVD_E_SHAPE, "expected a 2-D matrix")
        m, n = a.shape
        if m == 0 or n == 0:
            raise _lib.ShapeError(_lib.LSVD_E_SHAPE, "LieSvdSmall::solve_rectangular: empty input")
        k = min(m, n)
        h = handle or default_handle()
        u = np.zeros((m, k)); s = np.zeros(k); vt = np.zeros((k, n))
        LieSvdSmall.last_info = _call(h, _lib.lib().lsvd_small_solve_rectangular_f64, _ptr(a), m, n, _ptr(u),
                                      _ptr(s), _ptr(vt))
        return u, s, vt


def eigh_jacobi_full(p, handle: Handle | None = None):
    """Symmetric Jacobi eigensolver: returns (V, eig) with eig descending."""
    a = _f64(p)
    if a.ndim != 2 or a.shape[0] != a.shape[1]:
        raise _lib.ShapeError(_lib.LSVD_E_SHAPE, "eigh_jacobi_full: matrix must be square")
    n = a.shape[0]
    h = handle or default_handle()
    v = np.zeros((n, n)); e = np.zeros(n)
    if n == 0:
        return v, e
    eigh_jacobi_full.last_info = _call(h, _lib.lib().lsvd_eigh_jacobi_f64, _ptr(a), n, _ptr(v), _ptr(e))
    return v, e


class LieSvdMicro:
    """Fixed-schedule rotor SVD for N <= 4 (2x2 closed form)."""

    last_info: dict = {}

    @staticmethod
    def solve(mat, handle: Handle | None = None):
        a = _f64(mat)
        if a.ndim != 2 or a.shape[0] != a.shape[1]:
            raise _lib.ShapeError(_lib.LSVD_E_SHAPE, "LieSvdMicro: matrix must be square")
        n = a.shape[0]
        if n > MICRO_MAX_N:
            raise _lib.ShapeError(_lib.LSVD_E_SHAPE, "LieSvdMicro only supports N <= 4")
        h = handle or default_handle()
        u = np.zeros((n, n)); s = np.zeros(n); vt = np.zeros((n, n))
        if n == 0:
            return u, s, vt
        LieSvdMicro.last_info = _call(h, _lib.lib().lsvd_micro_solve_f64, _ptr(a), n, _ptr(u), _ptr(s), _ptr(vt))
        return u, s, vt


class LieSvd:
    """Public square entry (src/lie_svd.rs:49-51).

    The reference routes through LieSvdAdaptive: Micro for n <= 4, Small
    otherwise, unless its research triage picks PhaseFlow/CoreFlowTopo/Hybrid
    (src/lie_svd_adaptive.rs:103-119).  Those research routes are not part of
    this library, so the route here is Micro or Small.
    """

    @staticmethod
    def solve(mat, handle: Handle | None = None):
        a = _f64(mat)
        if a.ndim != 2 or a.shape[0] != a.shape[1]:
            raise _lib.ShapeError(_lib.LSVD_E_SHAPE, "LieSvd: matrix must be square")
        if a.shape[0] <= MICRO_MAX_N:
            return LieSvdMicro.solve(a, handle)
        return LieSvdSmall.solve(a, handle)


def run_small(a, handle: Handle | None = None):
    """solvers::run_small (src/lib.rs:56-60): (U, s, Vt, seconds)."""
    t0 = time.perf_counter()
    u, s, vt = LieSvdSmall.solve(a, handle)
    return u, s, vt, time.perf_counter() - t0


def svd_batched(mats, precision: int = _lib.LSVD_PREC_F64, handle: Handle | None = None, with_info=False):
    """Ragged batch in one plan execution (host arrays in, host arrays out).

    Square n <= 4 matrices take the LieSvdMicro schedule; the rest take
    LieSvdSmall.solve_rectangular.  Returns a list of (U, s, Vt)."""
    h = handle or default_handle()
    mats = [_f64(a) for a in mats]
    ms = [a.shape[0] for a in mats]
    ns = [a.shape[1] for a in mats]
    plan = BatchPlan(h, ms, ns, precision)
    L = plan.layout
    a_pool = np.concatenate([a.ravel() for a in mats]) if mats else np.zeros(1)
    u_pool = np.zeros(max(1, L["u_len"])); s_pool = np.zeros(max(1, L["s_len"]))
    vt_pool = np.zeros(max(1, L["vt_len"]))
    info = info_array(len(mats))
    plan.execute(a_pool, u_pool, s_pool, vt_pool, info)
    out = plan.split(u_pool, s_pool, vt_pool)
    plan.close()
    if with_info:
        return out, info_dicts(info, len(mats))
    return out


def digital_polish_batched(mats, u0s, vt0s, precision: int = _lib.LSVD_PREC_F64, handle: Handle | None = None,
                           with_info=False):
    """The digital-polish finishing step of the PhaseFlow, Analog and Hybrid
    routes (src/lie_svd_phaseflow.rs:643-658, src/lie_svd_analog.rs:170-184,
    src/lie_svd_hybrid.rs:108-116) for a batch of square matrices in one
    call: core = U0^T A V0, LieSvdSmall.solve(core), U = U0 Ur,
    Vt = Vrt Vt0.  Returns a list of (U, sigma, Vt)."""
    mats = [_f64(a) for a in mats]
    u0s = [_f64(a) for a in u0s]
    vt0s = [_f64(a) for a in vt0s]
    if not (len(mats) == len(u0s) == len(vt0s)):
        raise _lib.ShapeError(_lib.LSVD_E_SHAPE, "polish: as many U0 and Vt0 as matrices")
    for a, u0, vt0 in zip(mats, u0s, vt0s):
        n = a.shape[0]
        if a.ndim != 2 or a.shape != (n, n) or u0.shape != (n, n) or vt0.shape != (n, n) or n == 0:
            raise _lib.ShapeError(_lib.LSVD_E_SHAPE, "LieSvdSmall::solve: matrix must be square N x N")
    h = handle or default_handle()
    B = len(mats)
    if B == 0:
        return ([], []) if with_info else []
    n = np.array([a.shape[0] for a in mats], dtype=np.int32)
    mo = np.zeros(B, dtype=np.int64)
    mo[1:] = np.cumsum(n.astype(np.int64) ** 2)[:-1]
    so = np.zeros(B, dtype=np.int64)
    so[1:] = np.cumsum(n.astype(np.int64))[:-1]
    pa = np.concatenate([a.ravel() for a in mats])
    pu0 = np.concatenate([a.ravel() for a in u0s])
    pvt0 = np.concatenate([a.ravel() for a in vt0s])
    pu = np.zeros_like(pa); pvt = np.zeros_like(pa); ps = np.zeros(int(n.sum()))
    info = info_array(B)
    st = _lib.lib().lsvd_polish_batched_f64(h.raw, B, n.ctypes.data_as(_lib._i32p), mo.ctypes.data_as(_lib._i64p),
                                            so.ctypes.data_as(_lib._i64p), _ptr(pa), _ptr(pu0), _ptr(pvt0),
                                            int(precision), _ptr(pu), _ptr(ps), _ptr(pvt), _ptr(info))
    h.check(st)
    out = []
    for b in range(B):
        k = int(n[b]); o = int(mo[b]); q = int(so[b])
        out.append((pu[o:o + k * k].reshape(k, k).copy(), ps[q:q + k].copy(), pvt[o:o + k * k].reshape(k, k).copy()))
    return (out, info_dicts(info, B)) if with_info else out


def digital_polish(mat, u0, vt0, handle: Handle | None = None):
    """One matrix of digital_polish_batched: (U, sigma, Vt)."""
    return digital_polish_batched([mat], [u0], [vt0], handle=handle)[0]


__all__ = ["LieSvdSmall", "LieSvdMicro", "LieSvd", "eigh_jacobi_full", "run_small", "svd_batched",
           "digital_polish", "digital_polish_batched",
           "pack_offsets", "MICRO_MAX_N", "SMALL_MAX_N"]
