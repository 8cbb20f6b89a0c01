"""Newton-Schulz polar factor on the GPU (src/lie_svd_small.rs:158-183).

``newton_schulz_polar(a, precision)`` mirrors the reference's
``newton_schulz_polar`` (square) and extends it to tall matrices: it returns
``(Q, polar_ok)`` plus the diagnostics in ``last_info``.  ``precision`` picks
the arithmetic of the iterations (``LSVD_PREC_F64``: the reference
iteration; ``LSVD_PREC_F64_SCALED``: FP64 with scaled steps while small
singular values are still growing, same limit in fewer steps;
``LSVD_PREC_F32`` / ``LSVD_PREC_BF16``: MFMA products at that precision,
then FP64 refinement iterations).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .handle import Handle, _ptr, default_handle

last_info: dict = {}


def newton_schulz_polar(a, precision: int = _lib.LSVD_PREC_F64, handle: Handle | None = None):
    global last_info
    x = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if x.ndim != 2 or x.shape[0] < x.shape[1]:
        raise _lib.ShapeError(_lib.LSVD_E_SHAPE, "newton_schulz_polar: needs m >= n")
    m, n = x.shape
    h = handle or default_handle()
    q = np.zeros((m, n))
    info = _lib.LsvdPolarInfo()
    h.check(_lib.lib().lsvd_polar_f64(h.raw, _ptr(x), m, n, precision, _ptr(q), C.byref(info)))
    last_info = info.as_dict()
    return q, bool(info.polar_ok)


def polar_device(handle: Handle, a_dev, m: int, n: int, q_dev, precision: int) -> dict:
    """Device-pointer form (handle in LSVD_PTR_DEVICE mode); returns info."""
    info = _lib.LsvdPolarInfo()
    handle.check(_lib.lib().lsvd_polar_f64(handle.raw, _ptr(a_dev), m, n, precision, _ptr(q_dev), C.byref(info)))
    return info.as_dict()


def gemm_device(handle: Handle, trans_a: bool, m: int, n: int, k: int, a_dev, b_dev, c_dev, precision: int):
    """C = op(A) B on device buffers (FP64, or FP32 data with FP32/BF16 products)."""
    if precision in (_lib.LSVD_PREC_F64, _lib.LSVD_PREC_F64_SCALED):
        st = _lib.lib().lsvd_gemm_f64(handle.raw, int(trans_a), m, n, k, _ptr(a_dev), _ptr(b_dev), _ptr(c_dev))
    else:
        st = _lib.lib().lsvd_gemm_lowp(handle.raw, precision, int(trans_a), m, n, k, _ptr(a_dev), _ptr(b_dev),
                                       _ptr(c_dev))
    handle.check(st)
