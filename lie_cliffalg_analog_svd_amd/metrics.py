"""Residual metrics of the reference (src/lib.rs:133-194) plus the
FP64-floor-aware sigma errors used for parity reports.

These are host-side diagnostics over results already computed on the GPU
(the reference computes them on the CPU too); no solver work happens here.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class SvdMetrics:
    rel_recon: float
    orth_u: float
    orth_v: float
    sigma_max_rel: float | None = None
    sigma_tail_rel: float | None = None


def frobenius_norm(a) -> float:
    return float(np.sqrt(np.sum(np.asarray(a) ** 2)))


def compute(a, u, sigma, vt, sigma_ref=None) -> SvdMetrics:
    """metrics::compute (src/lib.rs:149-194).

    rel_recon = ||U S Vt - A||_F / max(||A||_F, 1e-300); orth_u = ||U^T U - I||_F;
    orth_v = ||Vt Vt^T - I||_F (from the returned Vt); sigma_max_rel and
    sigma_tail_rel (i >= floor(3n/4)) after sorting both descending.
    Identities are k x k, so thin rectangular factors are accepted too."""
    a = np.asarray(a, dtype=np.float64)
    u = np.asarray(u, dtype=np.float64)
    vt = np.asarray(vt, dtype=np.float64)
    s = np.asarray(sigma, dtype=np.float64)
    recon = (u * s[None, :]) @ vt
    rel = frobenius_norm(recon - a) / max(frobenius_norm(a), 1e-300)
    k = s.shape[0]
    orth_u = frobenius_norm(u.T @ u - np.eye(u.shape[1]))
    orth_v = frobenius_norm(vt @ vt.T - np.eye(vt.shape[0]))
    smr = stl = None
    if sigma_ref is not None:
        got = np.sort(s)[::-1]
        want = np.sort(np.asarray(sigma_ref, dtype=np.float64))[::-1]
        rel_e = np.abs(got - want) / np.maximum(np.abs(want), 1e-300)
        smr = float(rel_e.max()) if k else 0.0
        tail = (a.shape[0] * 3) // 4
        stl = float(rel_e[tail:].max()) if rel_e[tail:].size else 0.0
    return SvdMetrics(rel, orth_u, orth_v, smr, stl)


def sigma_errors(got, want, rel_floor: float = 1e-3):
    """Two-tier sigma error (the reading of src/lie_svd_benchmarks.rs:676-725):

    e_rel = max |s_i - r_i| / r_i over r_i >= rel_floor * r_max,
    e_abs = max |s_i - r_i| / r_max over all i."""
    g = np.sort(np.asarray(got, dtype=np.float64))[::-1]
    w = np.sort(np.asarray(want, dtype=np.float64))[::-1]
    if w.size == 0:
        return 0.0, 0.0
    smax = max(float(w[0]), 1e-300)
    d = np.abs(g - w)
    mask = w >= rel_floor * smax
    e_rel = float((d[mask] / w[mask]).max()) if mask.any() else 0.0
    return e_rel, float(d.max() / smax)


def subspace_angle_clusters(u_got, u_ref, sigma_ref, rtol: float = 1e-8):
    """sin of the largest principal angle between computed and reference
    singular subspaces, per cluster of (numerically) equal singular values,
    as ||(I - R R^T) G||_2 (accurate for tiny angles).  Returns a list of
    (start, stop, sine)."""
    s = np.asarray(sigma_ref, dtype=np.float64)
    smax = max(float(s.max()) if s.size else 0.0, 1e-300)
    out = []
    i = 0
    while i < s.size:
        j = i + 1
        while j < s.size and abs(s[j] - s[j - 1]) <= rtol * smax:
            j += 1
        a = np.asarray(u_got)[:, i:j]
        b = np.asarray(u_ref)[:, i:j]
        r = a - b @ (b.T @ a)
        out.append((i, j, float(np.linalg.norm(r, 2)) if r.size else 0.0))
        i = j
    return out


def compute_batched(handle, ms, ns, layout, a_pool, u_pool, s_pool, vt_pool, sigma_ref_pool=None):
    """metrics::compute (src/lib.rs:149-194) for every matrix of a batch on
    the device (lsvd_metrics_batched_f64): pools in the batch plan's packed
    layout (`layout` from pack_offsets), numpy arrays in host pointer mode
    or device tensors in device mode; `sigma_ref_pool` has the layout of
    s_pool.  Returns a structured array with fields rel_recon, orth_u,
    orth_v, sigma_max_rel, sigma_tail_rel (NaN without a reference)."""
    import ctypes as C

    from . import _lib
    from .handle import _as_i32p, _as_i64p, _i64, _ptr
    ms = np.ascontiguousarray(ms, dtype=np.int32)
    ns = np.ascontiguousarray(ns, dtype=np.int32)
    count = int(ms.shape[0])
    offs = [_i64(layout[k]) for k in ("a_off", "u_off", "s_off", "vt_off")]
    dt = np.dtype([(f, np.float64) for f in ("rel_recon", "orth_u", "orth_v", "sigma_max_rel", "sigma_tail_rel")])
    dev_mode = handle.pointer_mode == _lib.LSVD_PTR_DEVICE
    if dev_mode:
        import torch
        out_t = torch.empty((max(count, 1), 5), dtype=torch.float64, device=a_pool.device)
        out_ptr = _ptr(out_t)
    else:
        out = np.zeros(max(count, 1), dtype=dt)
        out_ptr = C.c_void_p(out.ctypes.data)
    handle.check(_lib.lib().lsvd_metrics_batched_f64(
        handle.raw, count, _as_i32p(ms), _as_i32p(ns), *[_as_i64p(o) for o in offs], _ptr(a_pool), _ptr(u_pool),
        _ptr(s_pool), _ptr(vt_pool), _ptr(sigma_ref_pool), out_ptr))
    if dev_mode:
        out = np.ascontiguousarray(out_t.cpu().numpy()).view(dt).reshape(-1)
    return out[:count]
