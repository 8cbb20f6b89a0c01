"""Device handle and batch plans over the C ABI.

``Handle`` wraps ``lsvd_handle_t``: one HIP stream, a pointer mode and an
autotune cache.  Host-mode calls take numpy arrays and are synchronous, like
the reference.  Device-mode calls take torch CUDA tensors (device memory)
and run asynchronously on the handle's stream, which can be bound to torch's
current stream.
"""
from __future__ import annotations

import ctypes as C
import threading

import numpy as np

from . import _lib
from ._lib import check, lib

_default = {}
_default_lock = threading.Lock()


def _ptr(a) -> C.c_void_p:
    """Raw address of a numpy array or a torch tensor."""
    if isinstance(a, np.ndarray):
        return C.c_void_p(a.ctypes.data)
    if hasattr(a, "data_ptr"):
        return C.c_void_p(a.data_ptr())
    if a is None:
        return C.c_void_p(0)
    raise TypeError(f"unsupported buffer type {type(a)}")


def _i32(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.int32))


def _i64(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.int64))


def _as_i32p(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _as_i64p(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


class Handle:
    """One lsvd_handle_t on `device`."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        st = lib().lsvd_create(device, C.byref(h))
        if st != _lib.LSVD_OK:
            raise _lib.LsvdError(st, f"lsvd_create(device={device}) failed (is a GPU visible?)")
        self._h = h
        self.device = device
        self.pointer_mode = _lib.LSVD_PTR_HOST

    @property
    def raw(self):
        return self._h

    def close(self):
        if self._h:
            lib().lsvd_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def set_pointer_mode(self, mode: int):
        check(self._h, lib().lsvd_set_pointer_mode(self._h, mode))
        self.pointer_mode = mode

    def set_solve_precision(self, precision: int):
        """Precision mode of the single-matrix solves (LSVD_PREC_F64, the
        reference iteration, by default; LSVD_PREC_F64_ONESIDED runs
        Householder QR + one-sided Jacobi)."""
        check(self._h, lib().lsvd_set_solve_precision(self._h, precision))

    def autotune(self, ms, ns, precision: int = _lib.LSVD_PREC_F64) -> dict:
        """lsvd_autotune_batch for a batch's shapes; returns the report
        (chosen plan knobs and the default / tuned plan times)."""
        ms = _i32(ms)
        ns = _i32(ns)
        check(self._h, lib().lsvd_autotune_batch(self._h, int(ms.shape[0]), _as_i32p(ms), _as_i32p(ns), precision))
        rep = _lib.LsvdTuneReport()
        check(self._h, lib().lsvd_autotune_report(self._h, C.byref(rep)))
        return rep.as_dict()

    def autotune_pools(self, ms, ns, layout, a_pool, u_pool, s_pool, vt_pool,
                       precision: int = _lib.LSVD_PREC_F64) -> dict:
        """lsvd_autotune_batch_pools: the plan knobs tuned on the caller's
        device pools (packed layout from pack_offsets); returns the report."""
        ms = _i32(ms)
        ns = _i32(ns)
        offs = [_i64(layout[k]) for k in ("a_off", "u_off", "s_off", "vt_off")]
        check(self._h, lib().lsvd_autotune_batch_pools(self._h, int(ms.shape[0]), _as_i32p(ms), _as_i32p(ns),
                                                         *[_as_i64p(o) for o in offs], precision, _ptr(a_pool),
                                                         _ptr(u_pool), _ptr(s_pool), _ptr(vt_pool)))
        rep = _lib.LsvdTuneReport()
        check(self._h, lib().lsvd_autotune_report(self._h, C.byref(rep)))
        return rep.as_dict()

    def workspace_info(self) -> tuple[int, int]:
        """(bytes, chunks) of the device workspace the handle keeps for its
        entry points' temporaries (lsvd_workspace_info)."""
        b = C.c_size_t()
        c = C.c_int()
        check(self._h, lib().lsvd_workspace_info(self._h, C.byref(b), C.byref(c)))
        return int(b.value), int(c.value)

    def set_stream(self, stream_ptr: int | None):
        check(self._h, lib().lsvd_set_stream(self._h, C.c_void_p(stream_ptr or 0)))

    def synchronize(self):
        check(self._h, lib().lsvd_synchronize(self._h))

    def check(self, status: int):
        check(self._h, status)


def default_handle(device: int = 0) -> Handle:
    """Process-wide host-mode handle used by the reference-style API."""
    with _default_lock:
        h = _default.get(device)
        if h is None:
            h = Handle(device)
            _default[device] = h
        return h


def pack_offsets(ms, ns, eigh: bool = False):
    """Packed pool offsets for row-major matrices of shapes m[b] x n[b]."""
    ms = _i64(ms)
    ns = _i64(ns)
    ks = np.minimum(ms, ns)
    a_sz = ms * ns
    u_sz = ns * ns if eigh else ms * ks
    vt_sz = np.zeros_like(ks) if eigh else ks * ns

    def offs(sz):
        o = np.zeros(len(sz), dtype=np.int64)
        if len(sz) > 1:
            o[1:] = np.cumsum(sz)[:-1]
        return o, int(sz.sum())

    (a_off, a_len), (u_off, u_len), (s_off, s_len), (vt_off, vt_len) = (
        offs(a_sz), offs(u_sz), offs(ks), offs(vt_sz))
    return dict(a_off=a_off, u_off=u_off, s_off=s_off, vt_off=vt_off,
                a_len=a_len, u_len=u_len, s_len=s_len, vt_len=vt_len, k=ks)


class BatchPlan:
    """lsvd_batch_plan_t: a sorted, bucketed ragged batch on concurrent lanes."""

    def __init__(self, handle: Handle, ms, ns, precision: int = _lib.LSVD_PREC_F64):
        self.handle = handle
        self.ms = _i32(ms)
        self.ns = _i32(ns)
        self.count = len(self.ms)
        self.layout = pack_offsets(self.ms, self.ns)
        self._offs = {k: _i64(self.layout[k]) for k in ("a_off", "u_off", "s_off", "vt_off")}
        p = C.c_void_p()
        st = lib().lsvd_batch_plan_create(
            handle.raw, self.count, _as_i32p(self.ms), _as_i32p(self.ns), _as_i64p(self._offs["a_off"]),
            _as_i64p(self._offs["u_off"]), _as_i64p(self._offs["s_off"]), _as_i64p(self._offs["vt_off"]),
            precision, C.byref(p))
        check(handle.raw, st)
        self._p = p

    def execute(self, a_pool, u_pool, s_pool, vt_pool, info=None):
        st = lib().lsvd_batch_plan_execute(self._p, _ptr(a_pool), _ptr(u_pool), _ptr(s_pool), _ptr(vt_pool),
                                           _ptr(info) if info is not None else C.c_void_p(0))
        check(self.handle.raw, st)

    def close(self):
        if self._p:
            lib().lsvd_batch_plan_destroy(self._p)
            self._p = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def describe(self):
        """Per-launch descriptors and the queue order (caller indices)."""
        nb = C.c_int(0)
        check(self.handle.raw, lib().lsvd_batch_plan_describe(self._p, None, 0, C.byref(nb), None))
        arr = (_lib.LsvdBucketDesc * max(1, nb.value))()
        order = np.zeros(max(1, self.count), dtype=np.int32)
        check(self.handle.raw, lib().lsvd_batch_plan_describe(self._p, arr, nb.value, C.byref(nb),
                                                               C.c_void_p(order.ctypes.data)))
        return [arr[i].as_dict() for i in range(nb.value)], order

    def time_buckets(self, a_pool, u_pool, s_pool, vt_pool):
        """Per-launch times (ms) of every bucket, each measured with HIP events
        on its lane while the other lanes run beside it (device pointers)."""
        descs, _ = self.describe()
        ms = (C.c_float * max(1, len(descs)))()
        check(self.handle.raw, lib().lsvd_batch_plan_time_buckets(
            self._p, _ptr(a_pool), _ptr(u_pool), _ptr(s_pool), _ptr(vt_pool), ms, len(descs)))
        return [float(ms[i]) for i in range(len(descs))]

    def phase_clocks(self, a_pool, u_pool, s_pool, vt_pool, clocks):
        """One plain run with per-matrix phase clocks (device int64 [count, 8])."""
        check(self.handle.raw, lib().lsvd_batch_plan_phase_clocks(
            self._p, _ptr(a_pool), _ptr(u_pool), _ptr(s_pool), _ptr(vt_pool), _ptr(clocks)))

    def split(self, u_pool, s_pool, vt_pool):
        """Views of the packed outputs: list of (U, s, Vt) numpy arrays."""
        L = self.layout
        out = []
        for b in range(self.count):
            m, n, k = int(self.ms[b]), int(self.ns[b]), int(L["k"][b])
            out.append((u_pool[L["u_off"][b]:L["u_off"][b] + m * k].reshape(m, k),
                        s_pool[L["s_off"][b]:L["s_off"][b] + k],
                        vt_pool[L["vt_off"][b]:L["vt_off"][b] + k * n].reshape(k, n)))
        return out


def info_array(count: int) -> np.ndarray:
    """Host buffer for `count` lsvd_info_t entries (8 int32 each)."""
    return np.zeros((max(count, 1), 8), dtype=np.int32)


INFO_FIELDS = [f for f, _ in _lib.LsvdInfo._fields_]


def info_dicts(arr: np.ndarray, count: int):
    return [dict(zip(INFO_FIELDS, map(int, arr[b]))) for b in range(count)]
