"""CPU FP64 oracle for the dense SVD path -- TEST INFRASTRUCTURE ONLY.

This package wraps ``liblsvd_oracle.so`` (built from ``lsvd_oracle.c`` by
``oracle/Makefile``), a C restatement of the reference Rust crate
skovnats/lie_cliffalg_analog_svd:

- ``src/lie_svd_small.rs:67-418``  (LieSvdSmall: NS polar, Jacobi, QR)
- ``src/lie_svd_micro.rs:15-214``  (LieSvdMicro: n <= 4)
- ``src/lie_svd_joint.rs:18-539``  (LieSvdJoint: Phase-JADE)
- ``src/lib.rs:149-194``           (metrics::compute)
- ``src/lie_svd_compiler.rs:144-268`` (HardwareSchedule: Givens export, JSON)
- ``src/lie_svd_phaseflow.rs:643-658`` (the digital-polish finishing step)
- ``src/lie_svd_adaptive.rs:59-392``, ``src/lie_svd_phasehealth.rs:100-345``,
  ``src/lie_svd_block4.rs:210-266`` (adaptive triage and router)

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import this package, and only as the checker or as
the timed CPU baseline.  The product package never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblsvd_oracle.so")
_lib = None


class Info(C.Structure):
    _fields_ = [
        ("polar_ok", C.c_int),
        ("ns_iters", C.c_int),
        ("jacobi_sweeps", C.c_int),
        ("repaired", C.c_int),
        ("micro_fallback", C.c_int),
        ("rotations", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class JointParams(C.Structure):
    _fields_ = [
        ("max_sweeps", C.c_int),
        ("tol", C.c_double),
        ("min_pair_scale", C.c_double),
        ("line_search_steps", C.c_int),
    ]


class JointTrace(C.Structure):
    _fields_ = [
        ("initial_offdiag", C.c_double),
        ("final_offdiag", C.c_double),
        ("sweeps", C.c_int),
        ("rotations", C.c_int),
        ("rejected_rotations", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PhaseEvent(C.Structure):
    """HardwarePhaseEvent (src/lie_svd_compiler.rs:27-41)."""
    _fields_ = [
        ("layer", C.c_int), ("i", C.c_int), ("j", C.c_int),
        ("phi_l", C.c_double), ("phi_r", C.c_double), ("theta", C.c_double),
        ("theta_l", C.c_double), ("theta_r", C.c_double),
        ("has_energy_before", C.c_int), ("has_energy_after", C.c_int),
        ("energy_before", C.c_double), ("energy_after", C.c_double),
        ("source", C.c_char_p), ("kind", C.c_char_p),
    ]


class Triage(C.Structure):
    """AdaptiveTriage (src/lie_svd_adaptive.rs:18-50) plus the route."""
    _fields_ = [("n", C.c_int), ("route", C.c_int)] + [(f, C.c_double) for f in (
        "offdiag_ratio", "diagonal_dominance", "row_cv", "col_cv", "row_col_mismatch", "symmetry_ratio",
        "transpose_torsion_ratio", "diagonal_entropy", "row_mass_entropy", "col_mass_entropy", "phase_mean_stress",
        "phase_max_twist", "phase_causal_disbalance", "phase_entropy_gap", "phase_torsion_energy",
        "phase_chirality_balance", "phase_entropy", "suspicious_score")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class AdaptiveParams(C.Structure):
    """LieSvdAdaptiveParams (src/lie_svd_adaptive.rs:59-78)."""
    _fields_ = [("small_max_n", C.c_int), ("coreflow_topo_max_n", C.c_int), ("phaseflow_max_n", C.c_int),
                ("nearly_diagonal_ratio", C.c_double), ("suspicious_threshold", C.c_double)]


ROUTES = ["Micro", "Small", "PhaseFlow", "CoreFlowTopo", "Hybrid"]


def build() -> None:
    """Compile the oracle with its Makefile (gcc, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        i64p = C.POINTER(C.c_int64)
        L.lso_small_solve.argtypes = [dp, C.c_int, dp, dp, dp, C.POINTER(Info)]
        L.lso_small_solve_rectangular.argtypes = [dp, C.c_int, C.c_int, dp, dp, dp, C.POINTER(Info)]
        L.lso_eigh_jacobi_full.argtypes = [dp, C.c_int, dp, dp, C.POINTER(Info)]
        L.lso_newton_schulz_polar.argtypes = [dp, C.c_int, C.c_int, dp, ip]
        L.lso_newton_schulz_polar.restype = C.c_int
        L.lso_gram_schmidt_orthonormal.argtypes = [dp, C.c_int, dp]
        L.lso_orthogonality_err.argtypes = [dp, C.c_int, C.c_int]
        L.lso_orthogonality_err.restype = C.c_double
        L.lso_qr_reduce.argtypes = [dp, C.c_int, C.c_int, dp, dp]
        L.lso_micro_solve.argtypes = [dp, C.c_int, dp, dp, dp, C.POINTER(Info)]
        L.lso_micro_solve.restype = C.c_int
        L.lso_micro_solve_capped.argtypes = [dp, C.c_int, C.c_int, dp, dp, dp, C.POINTER(Info)]
        L.lso_micro_solve_capped.restype = C.c_int
        L.lso_joint_params_for_n.argtypes = [C.c_int]
        L.lso_joint_params_for_n.restype = JointParams
        L.lso_joint_diag_sym.argtypes = [dp, C.c_int, C.c_int, JointParams, dp, dp, C.POINTER(JointTrace)]
        L.lso_joint_svd.argtypes = [dp, C.c_int, C.c_int, C.c_int, JointParams, dp, dp, dp, C.POINTER(JointTrace)]
        L.lso_layer_count.argtypes = [C.c_int]
        L.lso_layer_count.restype = C.c_int
        L.lso_layer_pairs.argtypes = [C.c_int, C.c_int, ip]
        L.lso_layer_pairs.restype = C.c_int
        L.lso_metrics.argtypes = [dp, C.c_int, dp, dp, dp, dp, dp]
        L.lso_svd_batch.argtypes = [C.c_int, ip, ip, i64p, dp, i64p, dp, i64p, dp, i64p, dp, C.POINTER(Info),
                                    C.c_int]
        L.lso_joint_batch.argtypes = [C.c_int, C.c_int, C.c_int, dp, JointParams, dp, dp,
                                      C.POINTER(JointTrace), C.c_int]
        L.lso_ns_iterations.argtypes = [dp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.lso_svd_loss_grad.argtypes = [dp, C.c_int, C.c_int, dp, dp, dp]
        L.lso_svd_loss_grad.restype = C.c_double
        L.lso_adaptive_params_default.restype = AdaptiveParams
        L.lso_analyze_matrix.argtypes = [dp, C.c_int, C.POINTER(Triage)]
        for fn in ("lso_choose_route", "lso_should_use_phaseflow", "lso_should_use_coreflow_topo",
                   "lso_is_balanced_degenerate"):
            getattr(L, fn).argtypes = [C.POINTER(Triage), C.POINTER(AdaptiveParams)]
            getattr(L, fn).restype = C.c_int
        L.lso_polish.argtypes = [dp, dp, dp, C.c_int, dp, dp, dp, C.POINTER(Info)]
        L.lso_hosvd3.argtypes = [dp, C.c_int, C.c_int, C.c_int, dp, dp, dp, dp, dp]
        L.lso_hosvd3.restype = None
        L.lso_rotate_tensor3.argtypes = [dp, C.c_int, C.c_int, C.c_int, dp, C.c_int, dp, C.c_int, dp, C.c_int, dp]
        L.lso_rotate_tensor3.restype = None
        L.lso_givens_schedule.argtypes = [dp, C.c_int, ip, ip, dp, dp, dp, dp]
        L.lso_givens_schedule.restype = C.c_int
        L.lso_schedule_json.argtypes = [C.c_char_p, C.c_int, C.c_int, C.POINTER(PhaseEvent), C.c_int, dp,
                                        C.c_char_p, C.c_size_t]
        L.lso_schedule_json.restype = C.c_size_t
        L.lso_bss_whiten.argtypes = [dp, C.c_int, C.c_int, ip, C.c_int, C.c_double, dp, dp, dp, dp, dp]
        L.lso_bss_whiten.restype = C.c_int
        L.lso_bss_unmix.argtypes = [dp, dp, dp, C.c_int, C.c_int, dp, dp, dp]
        L.lso_bss_unmix.restype = None
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _c(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def small_solve(a):
    """LieSvdSmall::solve (src/lie_svd_small.rs:296-323) -> (U, s, Vt, info)."""
    a = _c(a)
    n = a.shape[0]
    assert a.shape == (n, n), "LieSvdSmall::solve: matrix must be square N x N"
    u = np.zeros((n, n)); s = np.zeros(n); vt = np.zeros((n, n))
    info = Info()
    lib().lso_small_solve(_dp(a), n, _dp(u), _dp(s), _dp(vt), C.byref(info))
    return u, s, vt, info.as_dict()


def small_solve_rectangular(a):
    """LieSvdSmall::solve_rectangular (src/lie_svd_small.rs:340-360)."""
    a = _c(a)
    m, n = a.shape
    assert m > 0 and n > 0, "LieSvdSmall::solve_rectangular: empty input"
    k = min(m, n)
    u = np.zeros((m, k)); s = np.zeros(k); vt = np.zeros((k, n))
    info = Info()
    lib().lso_small_solve_rectangular(_dp(a), m, n, _dp(u), _dp(s), _dp(vt), C.byref(info))
    return u, s, vt, info.as_dict()


def eigh_jacobi_full(p):
    """eigh_jacobi_full (src/lie_svd_small.rs:212-285) -> (V, eig, info)."""
    p = _c(p)
    n = p.shape[0]
    v = np.zeros((n, n)); e = np.zeros(n)
    info = Info()
    lib().lso_eigh_jacobi_full(_dp(p), n, _dp(v), _dp(e), C.byref(info))
    return v, e, info.as_dict()


def newton_schulz_polar(a):
    """newton_schulz_polar (src/lie_svd_small.rs:158-183) -> (Q, ok, iters)."""
    a = _c(a)
    m, n = a.shape
    q = np.zeros((m, n))
    it = C.c_int(0)
    ok = lib().lso_newton_schulz_polar(_dp(a), m, n, _dp(q), C.byref(it))
    return q, bool(ok), it.value


def gram_schmidt_orthonormal(a):
    a = _c(a)
    n = a.shape[0]
    out = np.zeros((n, n))
    lib().lso_gram_schmidt_orthonormal(_dp(a), n, _dp(out))
    return out


def qr_reduce(a):
    """qr_reduce (src/lie_svd_small.rs:373-418) -> (Q, R)."""
    a = _c(a)
    m, d = a.shape
    q = np.zeros((m, d)); r = np.zeros((d, d))
    lib().lso_qr_reduce(_dp(a), m, d, _dp(q), _dp(r))
    return q, r


def micro_solve(a, sweep_cap: int = 0):
    """LieSvdMicro::solve (src/lie_svd_micro.rs:21-52) -> (U, s, Vt, info).
    sweep_cap > 0 lowers the 24-sweep limit of n = 3, 4 (a test hook that
    forces the escalation to LieSvdSmall, :90-92)."""
    a = _c(a)
    n = a.shape[0]
    assert a.shape == (n, n) and n <= 4
    u = np.zeros((n, n)); s = np.zeros(n); vt = np.zeros((n, n))
    info = Info()
    rc = lib().lso_micro_solve_capped(_dp(a), n, int(sweep_cap), _dp(u), _dp(s), _dp(vt), C.byref(info))
    assert rc == 0
    return u, s, vt, info.as_dict()


def joint_params_for_n(n: int) -> JointParams:
    return lib().lso_joint_params_for_n(n)


def joint_diag_sym(mats, params: JointParams | None = None):
    """LieSvdJoint::diagonalize_symmetric_with_params (src/lie_svd_joint.rs:71-158).

    mats: (K, n, n).  Returns (V, diagonals (K, n), trace dict)."""
    mats = _c(mats)
    k, n, n2 = mats.shape
    assert n == n2
    if params is None:
        params = joint_params_for_n(n)
    v = np.zeros((n, n)); d = np.zeros((k, n))
    tr = JointTrace()
    lib().lso_joint_diag_sym(_dp(mats), k, n, params, _dp(v), _dp(d), C.byref(tr))
    return v, d, tr.as_dict()


def joint_svd(mats, params: JointParams | None = None):
    """LieSvdJoint::joint_svd_with_params (src/lie_svd_joint.rs:170-262).

    mats: (K, rows, cols).  Returns (U (rows, rows), sigmas (K, min), Vt
    (cols, cols), trace dict); the reference's panics are assertions."""
    mats = _c(mats)
    assert mats.ndim == 3 and mats.shape[0] > 0, "LieSvdJoint: empty matrix set"
    k, rows, cols = mats.shape
    assert rows > 0 and cols > 0, "LieSvdJoint: empty matrix"
    kk = min(rows, cols)
    if params is None:
        params = joint_params_for_n(kk)
    u = np.zeros((rows, rows)); sg = np.zeros((k, kk)); vt = np.zeros((cols, cols))
    tr = JointTrace()
    lib().lso_joint_svd(_dp(mats), k, rows, cols, params, _dp(u), _dp(sg), _dp(vt), C.byref(tr))
    return u, sg, vt, tr.as_dict()


def layer_pairs(n: int, layer: int):
    """layer_pairs (src/lie_svd_joint.rs:524-539)."""
    buf = (C.c_int * (n + 2))()
    cnt = lib().lso_layer_pairs(n, layer, buf)
    return [(buf[2 * i], buf[2 * i + 1]) for i in range(cnt)]


def layer_count(n: int) -> int:
    return lib().lso_layer_count(n)


def metrics(a, u, s, vt, sigma_ref=None):
    """metrics::compute (src/lib.rs:149-194) -> dict."""
    a, u, s, vt = _c(a), _c(u), _c(s), _c(vt)
    n = a.shape[0]
    out = np.zeros(5)
    sr = _dp(_c(sigma_ref)) if sigma_ref is not None else None
    lib().lso_metrics(_dp(a), n, _dp(u), _dp(s), _dp(vt), sr, _dp(out))
    res = {"rel_recon": out[0], "orth_u": out[1], "orth_v": out[2]}
    if sigma_ref is not None:
        res["sigma_max_rel"] = out[3]
        res["sigma_tail_rel"] = out[4]
    return res


def svd_batch(mats, threads: int = 0, with_info: bool = False):
    """Ragged batch through micro (square n <= 4) or solve_rectangular, one
    matrix per OpenMP thread.  Returns a list of (U, s, Vt), plus the
    per-matrix info dicts when with_info."""
    count = len(mats)
    ms = np.array([a.shape[0] for a in mats], dtype=np.int32)
    ns = np.array([a.shape[1] for a in mats], dtype=np.int32)
    ks = np.minimum(ms, ns)
    a_sz = ms.astype(np.int64) * ns
    u_sz = ms.astype(np.int64) * ks
    vt_sz = ks.astype(np.int64) * ns
    offs = lambda sz: np.concatenate([[0], np.cumsum(sz)[:-1]]).astype(np.int64)
    a_off, u_off, s_off, vt_off = offs(a_sz), offs(u_sz), offs(ks.astype(np.int64)), offs(vt_sz)
    a_pool = np.concatenate([_c(a).ravel() for a in mats]) if count else np.zeros(1)
    u_pool = np.zeros(max(1, int(u_sz.sum()))); s_pool = np.zeros(max(1, int(ks.sum())))
    vt_pool = np.zeros(max(1, int(vt_sz.sum())))
    ip = lambda x: x.ctypes.data_as(C.POINTER(C.c_int))
    lp = lambda x: x.ctypes.data_as(C.POINTER(C.c_int64))
    infos = (Info * max(1, count))()
    lib().lso_svd_batch(count, ip(ms), ip(ns), lp(a_off), _dp(a_pool), lp(u_off), _dp(u_pool),
                        lp(s_off), _dp(s_pool), lp(vt_off), _dp(vt_pool), infos, threads)
    out = []
    for b in range(count):
        m, n, k = int(ms[b]), int(ns[b]), int(ks[b])
        out.append((u_pool[u_off[b]:u_off[b] + m * k].reshape(m, k),
                    s_pool[s_off[b]:s_off[b] + k].copy(),
                    vt_pool[vt_off[b]:vt_off[b] + k * n].reshape(k, n)))
    if with_info:
        return out, [infos[b].as_dict() for b in range(count)]
    return out


def joint_batch(families, params: JointParams | None = None, threads: int = 0):
    """families: (F, K, n, n) -> (V (F,n,n), diag (F,K,n), traces)."""
    families = _c(families)
    f, k, n, _ = families.shape
    if params is None:
        params = joint_params_for_n(n)
    v = np.zeros((f, n, n)); d = np.zeros((f, k, n))
    traces = (JointTrace * f)()
    lib().lso_joint_batch(f, k, n, _dp(families), params, _dp(v), _dp(d), traces, threads)
    return v, d, [t.as_dict() for t in traces]


def ns_iterations(x, iters: int, threads: int = 0):
    """Runs exactly `iters` NS iterations in place (CPU baseline sample)."""
    assert x.dtype == np.float64 and x.flags.c_contiguous
    m, n = x.shape
    lib().lso_ns_iterations(_dp(x), m, n, iters, threads)
    return x


def svd_loss_grad(a, w, g):
    """FP64 analytic gradient of L = sum_i w_i s_i + <G, U Vt> (no reference
    counterpart; closed form in lsvd_oracle.c).  Returns (loss, dL/dA)."""
    a, w, g = _c(a), _c(w), _c(g)
    m, n = a.shape
    grad = np.zeros((m, n))
    loss = lib().lso_svd_loss_grad(_dp(a), m, n, _dp(w), _dp(g), _dp(grad))
    return loss, grad


def givens_schedule(v):
    """HardwareSchedule::from_orthogonal_matrix (src/lie_svd_compiler.rs:
    144-197) -> (events (n, 5): columns i, j, c, s, theta; diagonal signs).
    Event e's layer is e."""
    v = _c(v)
    d = v.shape[0]
    assert v.shape == (d, d), f"HardwareSchedule::from_orthogonal_matrix: matrix must be square, got {v.shape}"
    cap = max(d * (d - 1) // 2, 1)
    ei = np.zeros(cap, dtype=np.int32); ej = np.zeros(cap, dtype=np.int32)
    ec = np.zeros(cap); es = np.zeros(cap); et = np.zeros(cap); sg = np.zeros(max(d, 1))
    ip = C.POINTER(C.c_int)
    n = lib().lso_givens_schedule(_dp(v), d, ei.ctypes.data_as(ip), ej.ctypes.data_as(ip), _dp(ec), _dp(es),
                                  _dp(et), _dp(sg))
    ev = np.zeros((n, 5))
    ev[:, 0] = ei[:n]; ev[:, 1] = ej[:n]; ev[:, 2] = ec[:n]; ev[:, 3] = es[:n]; ev[:, 4] = et[:n]
    return ev, sg[:d].copy()


def schedule_json(target: str, channels: int, events, signs=()) -> str:
    """HardwareSchedule::to_json_string (src/lie_svd_compiler.rs:213-268).
    events: dicts with the HardwarePhaseEvent fields (energy_before /
    energy_after optional)."""
    arr = (PhaseEvent * max(len(events), 1))()
    keep = []
    for k, e in enumerate(events):
        src = e["source"].encode(); kind = e["kind"].encode()
        keep += [src, kind]
        arr[k] = PhaseEvent(int(e["layer"]), int(e["i"]), int(e["j"]), float(e["phi_l"]), float(e["phi_r"]),
                            float(e["theta"]), float(e["theta_l"]), float(e["theta_r"]),
                            int(e.get("energy_before") is not None), int(e.get("energy_after") is not None),
                            float(e.get("energy_before") or 0.0), float(e.get("energy_after") or 0.0), src, kind)
    sg = _c(np.asarray(signs, dtype=np.float64).reshape(-1))
    if sg.size == 0:
        sg = np.zeros(1)
        ns = 0
    else:
        ns = sg.size
    L = lib()
    need = L.lso_schedule_json(target.encode(), channels, len(events), arr, ns, _dp(sg), None, 0)
    buf = C.create_string_buffer(need + 1)
    L.lso_schedule_json(target.encode(), channels, len(events), arr, ns, _dp(sg), buf, need + 1)
    return buf.value.decode()


def givens_events(ev):
    """The HardwarePhaseEvent dicts from_orthogonal_matrix records
    (src/lie_svd_compiler.rs:172-186) for givens_schedule's events."""
    return [{"layer": k, "i": int(r[0]), "j": int(r[1]), "phi_l": 0.0, "phi_r": 0.0, "theta": r[4],
             "theta_l": r[4], "theta_r": 0.0, "source": "orthogonal_givens", "kind": "givens_elimination"}
            for k, r in enumerate(ev)]


def polish(a, u0, vt0):
    """solve_with_digital_polish's finishing step (src/lie_svd_phaseflow.rs:
    651-657): core = U0^T A Vt0^T, LieSvdSmall::solve(core), U = U0 Ur,
    Vt = Vrt Vt0 -> (U, s, Vt, info of the core's solve)."""
    a, u0, vt0 = _c(a), _c(u0), _c(vt0)
    n = a.shape[0]
    assert a.shape == (n, n) and u0.shape == (n, n) and vt0.shape == (n, n), "polish: square n x n inputs"
    u = np.zeros((n, n)); s = np.zeros(n); vt = np.zeros((n, n))
    info = Info()
    lib().lso_polish(_dp(a), _dp(u0), _dp(vt0), n, _dp(u), _dp(s), _dp(vt), C.byref(info))
    return u, s, vt, info.as_dict()


def adaptive_params_default() -> AdaptiveParams:
    return lib().lso_adaptive_params_default()


def analyze_matrix(a) -> dict:
    """analyze_matrix (src/lie_svd_adaptive.rs:186-274) -> triage dict."""
    a = _c(a)
    n = a.shape[0]
    assert a.shape == (n, n), "adaptive SVD expects square matrices"
    t = Triage()
    lib().lso_analyze_matrix(_dp(a), n, C.byref(t))
    return t.as_dict()


def _triage(d: dict) -> Triage:
    t = Triage()
    for k, _ in Triage._fields_:
        if k in d:
            setattr(t, k, d[k])
    return t


def choose_route(a, params: AdaptiveParams | None = None) -> tuple:
    """choose_route (src/lie_svd_adaptive.rs:103-119) -> (route name, triage)."""
    d = analyze_matrix(a)
    p = params or adaptive_params_default()
    r = lib().lso_choose_route(C.byref(_triage(d)), C.byref(p))
    d["route"] = r
    return ROUTES[r], d


def route_predicate(name: str, triage: dict, params: AdaptiveParams | None = None) -> bool:
    """should_use_phaseflow / should_use_coreflow_topo /
    is_balanced_degenerate_case on a triage dict."""
    p = params or adaptive_params_default()
    fn = {"phaseflow": "lso_should_use_phaseflow", "coreflow_topo": "lso_should_use_coreflow_topo",
          "balanced_degenerate": "lso_is_balanced_degenerate"}[name]
    return bool(getattr(lib(), fn)(C.byref(_triage(triage)), C.byref(p)))


def hosvd3(x):
    """LieSvdTensor::hosvd3 (src/lie_svd_tensor.rs:31-62).  Returns (core, U1,
    U2, U3, trace dict)."""
    x = _c(x)
    n1, n2, n3 = x.shape
    core = np.zeros_like(x)
    us = [np.zeros((n, n)) for n in (n1, n2, n3)]
    tr = np.zeros(3)
    lib().lso_hosvd3(_dp(x), n1, n2, n3, _dp(core), _dp(us[0]), _dp(us[1]), _dp(us[2]), _dp(tr))
    return core, us[0], us[1], us[2], {"initial_offdiag": tr[0], "final_offdiag": tr[1],
                                       "superdiag_mass_ratio": tr[2]}


def rotate_tensor3(x, q1, q2, q3):
    """rotate_tensor3 (src/lie_svd_tensor.rs:121-145): q_m is r_m x n_m."""
    x = _c(x)
    q1, q2, q3 = _c(q1), _c(q2), _c(q3)
    n1, n2, n3 = x.shape
    out = np.zeros((q1.shape[0], q2.shape[0], q3.shape[0]))
    lib().lso_rotate_tensor3(_dp(x), n1, n2, n3, _dp(q1), q1.shape[0], _dp(q2), q2.shape[0], _dp(q3), q3.shape[0],
                             _dp(out))
    return out


BSS_LAGS = (1, 2, 3, 5)  # PhaseBssParams::for_channels (src/lie_svd_bss.rs:20-26)
BSS_EPS = 1e-10


def bss_whiten(x, lags=BSS_LAGS, eps: float = BSS_EPS):
    """The first half of LieSvdBss::separate (src/lie_svd_bss.rs:58-75,
    162-197) for one channels x samples observation block.  Returns
    (whitening W, whitened, lagged family (K, ch, ch), sigma, floor)."""
    x = _c(x)
    ch, samples = x.shape
    lg = np.ascontiguousarray(np.asarray(lags, dtype=np.int32))
    w = np.zeros((ch, ch)); z = np.zeros((ch, samples)); fam = np.zeros((max(len(lg), 1), ch, ch))
    sg = np.zeros(ch); fl = np.zeros(1)
    k = lib().lso_bss_whiten(_dp(x), ch, samples, lg.ctypes.data_as(C.POINTER(C.c_int)), len(lg), eps, _dp(w),
                             _dp(z), _dp(fam), _dp(sg), _dp(fl))
    return w, z, fam[:k].copy(), sg, float(fl[0])


def bss_unmix(basis, whitening, whitened):
    """The second half (src/lie_svd_bss.rs:81-86, 104-126): (unmixing,
    separated, channel coherence)."""
    v, w, z = _c(basis), _c(whitening), _c(whitened)
    ch, samples = z.shape
    un = np.zeros((ch, ch)); sep = np.zeros((ch, samples)); coh = np.zeros(ch)
    lib().lso_bss_unmix(_dp(v), _dp(w), _dp(z), ch, samples, _dp(un), _dp(sep), _dp(coh))
    return un, sep, coh


def bss_separate(x, lags=BSS_LAGS, eps: float = BSS_EPS):
    """LieSvdBss::separate (src/lie_svd_bss.rs:58-101): whiten, Phase-JADE on
    the lagged family with for_n(channels), unmix.  Returns a dict."""
    w, z, fam, sg, fl = bss_whiten(x, lags, eps)
    v, _d, tr = joint_diag_sym(fam, joint_params_for_n(z.shape[0]))
    un, sep, coh = bss_unmix(v, w, z)
    return {"unmixing": un, "separated": sep, "channel_coherence": coh, "whitening": w, "whitened": z,
            "lagged": fam, "basis": v, "sigma": sg, "covariance_floor": fl, "joint": tr,
            "mean_channel_coherence": float(coh.sum() / max(len(coh), 1))}
