"""Reverse-mode adjoints of the SVD and the polar factor on the GPU.

No reference counterpart (the reference has no derivatives).  ``svd_vjp``
takes cotangents of (U, s, Vt) of the thin SVD of a batch of equally sized
m x n matrices (m >= n) and returns dL/dA; ``polar_vjp`` does the same for
the polar factor Q = U Vt.  When the forward factors are not passed, the
library recomputes them on the GPU instead of storing them.  ``guard`` sets
the broadening of the 1/(s_j - s_i) factor: eps = guard * s_max.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .handle import Handle, _ptr, default_handle


def _batch(x, shape_tail, name):
    if x is None:
        return None
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    if x.shape[1:] != shape_tail:
        raise _lib.ShapeError(_lib.LSVD_E_SHAPE, f"{name}: expected (*, {shape_tail})")
    return x


def svd_vjp(a, gu=None, gs=None, gvt=None, forward=None, guard: float = 1e-12, handle: Handle | None = None):
    """a: (B, m, n); gu (B, m, n), gs (B, n), gvt (B, n, n) may be None;
    forward: optional (U, s, Vt) batches.  Returns dL/dA (B, m, n)."""
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    single = a.ndim == 2
    if single:
        a = a[None]
    B, m, n = a.shape
    if m < n:
        raise _lib.ShapeError(_lib.LSVD_E_SHAPE, "svd_vjp: needs m >= n")

    def lift(x, nd):  # accept unbatched arrays for a single matrix
        return None if x is None else (np.asarray(x)[None] if np.ndim(x) == nd else x)

    gu = _batch(lift(gu, 2), (m, n), "gu")
    gs = _batch(lift(gs, 1), (n,), "gs")
    gvt = _batch(lift(gvt, 2), (n, n), "gvt")
    u = s = vt = None
    if forward is not None:
        u = _batch(lift(forward[0], 2), (m, n), "U")
        s = _batch(lift(forward[1], 1), (n,), "s")
        vt = _batch(lift(forward[2], 2), (n, n), "Vt")
    h = handle or default_handle()
    grad = np.zeros((B, m, n))
    st = _lib.lib().lsvd_svd_vjp_f64(h.raw, B, m, n, _ptr(a), _ptr(u), _ptr(s), _ptr(vt), _ptr(gu), _ptr(gs),
                                     _ptr(gvt), float(guard), _ptr(grad))
    h.check(st)
    return grad


def polar_vjp(a, qbar, guard: float = 1e-12, handle: Handle | None = None):
    """dL/dA for a loss of the polar factor Q = U Vt, given qbar = dL/dQ."""
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if a.ndim == 2:
        a = a[None]
    B, m, n = a.shape
    qbar = _batch(qbar if np.ndim(qbar) == 3 else np.asarray(qbar)[None], (m, n), "qbar")
    h = handle or default_handle()
    grad = np.zeros((B, m, n))
    h.check(_lib.lib().lsvd_polar_vjp_f64(h.raw, B, m, n, _ptr(a), _ptr(qbar), float(guard), _ptr(grad)))
    return grad


def svd_loss_grad(a, w, g, guard: float = 1e-12, handle: Handle | None = None):
    """dL/dA of L = sum_i w_i s_i + <G, U Vt> for a batch: the sigma part
    through svd_vjp (gs = w) and the polar part through polar_vjp (qbar = G)."""
    return svd_vjp(a, gs=w, guard=guard, handle=handle) + polar_vjp(a, g, guard=guard, handle=handle)
