"""ctypes binding of ``liblsvd_amd.so`` (the C ABI in include/lie_svd_amd.h).

The library is the product: there is no CPU fallback.  Loading fails with a
clear error when the shared library has not been built, and every call fails
when no GPU is present.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblsvd_amd.so")

LSVD_OK = 0
LSVD_E_SHAPE = 1
LSVD_E_HIP = 2
LSVD_E_NONFINITE = 3
LSVD_E_NOCONV = 4
LSVD_E_ARG = 5
LSVD_E_NOMEM = 6
LSVD_E_LIMIT = 7

LSVD_PTR_HOST = 0
LSVD_PTR_DEVICE = 1

LSVD_PREC_F64 = 0
LSVD_PREC_F32 = 1
LSVD_PREC_BF16 = 2
LSVD_PREC_F64_SCALED = 3
LSVD_PREC_F64_ONESIDED = 4

LSVD_TARGET_MZI_MESH = 0
LSVD_TARGET_FPGA_ROTOR_MESH = 1


class LsvdInfo(C.Structure):
    _fields_ = [
        ("polar_ok", C.c_int32),
        ("ns_iters", C.c_int32),
        ("jacobi_sweeps", C.c_int32),
        ("rotations", C.c_int32),
        ("repaired", C.c_int32),
        ("micro_fallback", C.c_int32),
        ("path", C.c_int32),
        ("refine_iters", C.c_int32),
    ]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class LsvdMetrics(C.Structure):
    _fields_ = [
        ("rel_recon", C.c_double),
        ("orth_u", C.c_double),
        ("orth_v", C.c_double),
        ("sigma_max_rel", C.c_double),
        ("sigma_tail_rel", C.c_double),
    ]


class LsvdTuneReport(C.Structure):
    _fields_ = [
        ("multi_min_np", C.c_int32),
        ("max_lanes", C.c_int32),
        ("wide_mask", C.c_int32),
        ("evaluations", C.c_int32),
        ("default_ms", C.c_double),
        ("tuned_ms", C.c_double),
        ("os_max_np", C.c_int32),
        ("lead_np", C.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class LsvdJointParams(C.Structure):
    _fields_ = [
        ("max_sweeps", C.c_int32),
        ("tol", C.c_double),
        ("min_pair_scale", C.c_double),
        ("line_search_steps", C.c_int32),
    ]


class LsvdJointTrace(C.Structure):
    _fields_ = [
        ("initial_offdiag", C.c_double),
        ("final_offdiag", C.c_double),
        ("sweeps", C.c_int32),
        ("rotations", C.c_int32),
        ("rejected_rotations", C.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class LsvdTensorTrace(C.Structure):
    _fields_ = [
        ("initial_offdiag", C.c_double),
        ("final_offdiag", C.c_double),
        ("superdiag_mass_ratio", C.c_double),
    ]


class LsvdBssTrace(C.Structure):
    _fields_ = [
        ("channels", C.c_int32),
        ("samples", C.c_int32),
        ("covariance_floor", C.c_double),
        ("joint_initial_offdiag", C.c_double),
        ("joint_final_offdiag", C.c_double),
        ("joint_rotations", C.c_int32),
        ("mean_channel_coherence", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class LsvdPolarInfo(C.Structure):
    _fields_ = [
        ("polar_ok", C.c_int32),
        ("lowp_iters", C.c_int32),
        ("f64_iters", C.c_int32),
        ("reserved", C.c_int32),
        ("orth_err", C.c_double),
        ("last_delta", C.c_double),
        ("skew_err", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class LsvdPhaseEvent(C.Structure):
    """lsvd_phase_event_t (HardwarePhaseEvent, src/lie_svd_compiler.rs:27-41)."""
    _fields_ = [
        ("layer", C.c_int32), ("i", C.c_int32), ("j", C.c_int32), ("flags", C.c_int32),
        ("phi_l", C.c_double), ("phi_r", C.c_double), ("theta", C.c_double),
        ("theta_l", C.c_double), ("theta_r", C.c_double),
        ("energy_before", C.c_double), ("energy_after", C.c_double),
        ("source", C.c_int32), ("kind", C.c_int32),
    ]


TRIAGE_FIELDS = ["offdiag_ratio", "diagonal_dominance", "row_cv", "col_cv", "row_col_mismatch", "symmetry_ratio",
                 "transpose_torsion_ratio", "diagonal_entropy", "row_mass_entropy", "col_mass_entropy",
                 "phase_mean_stress", "phase_max_twist", "phase_causal_disbalance", "phase_entropy_gap",
                 "phase_torsion_energy", "phase_chirality_balance", "phase_entropy", "suspicious_score"]


class LsvdTriage(C.Structure):
    """lsvd_triage_t (AdaptiveTriage, src/lie_svd_adaptive.rs:18-50, and the route)."""
    _fields_ = [("n", C.c_int32), ("route", C.c_int32)] + [(f, C.c_double) for f in TRIAGE_FIELDS]


class LsvdAdaptiveParams(C.Structure):
    """lsvd_adaptive_params_t (LieSvdAdaptiveParams, src/lie_svd_adaptive.rs:59-78)."""
    _fields_ = [("small_max_n", C.c_int32), ("coreflow_topo_max_n", C.c_int32), ("phaseflow_max_n", C.c_int32),
                ("reserved", C.c_int32), ("nearly_diagonal_ratio", C.c_double), ("suspicious_threshold", C.c_double)]


class LsvdBucketDesc(C.Structure):
    _fields_ = [
        ("first", C.c_int32),
        ("count", C.c_int32),
        ("max_n", C.c_int32),
        ("variant", C.c_int32),
        ("grid", C.c_int32),
        ("threads", C.c_int32),
        ("lds_bytes", C.c_int64),
    ]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class LsvdError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"lsvd error {status} ({status_string(status)}): {msg}")
        self.status = status


class ShapeError(LsvdError, AssertionError):
    """The reference panics (assert!) on shape misuse; this mirrors it."""


_lib = None

_dp = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_vp = C.c_void_p

# name -> (restype, argtypes); symbols declared in include/lie_svd_amd.h
SIGNATURES = {
    "lsvd_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "lsvd_destroy": (C.c_int, [_vp]),
    "lsvd_set_stream": (C.c_int, [_vp, _vp]),
    "lsvd_get_stream": (C.c_void_p, [_vp]),
    "lsvd_set_pointer_mode": (C.c_int, [_vp, C.c_int]),
    "lsvd_set_solve_precision": (C.c_int, [_vp, C.c_int]),
    "lsvd_synchronize": (C.c_int, [_vp]),
    "lsvd_status_string": (C.c_char_p, [C.c_int]),
    "lsvd_last_error": (C.c_char_p, [_vp]),
    "lsvd_version": (C.c_char_p, []),
    "lsvd_small_solve_f64": (C.c_int, [_vp, _vp, C.c_int, _vp, _vp, _vp, _vp]),
    "lsvd_small_solve_rectangular_f64": (C.c_int, [_vp, _vp, C.c_int, C.c_int, _vp, _vp, _vp, _vp]),
    "lsvd_small_solve_rectangular_f64_noalloc": (C.c_int, [_vp, _vp, C.c_int, C.c_int, _vp, _vp, _vp, _vp]),
    "lsvd_eigh_jacobi_f64": (C.c_int, [_vp, _vp, C.c_int, _vp, _vp, _vp]),
    "lsvd_micro_solve_f64": (C.c_int, [_vp, _vp, C.c_int, _vp, _vp, _vp, _vp]),
    "lsvd_batch_plan_create": (C.c_int, [_vp, C.c_int, _i32p, _i32p, _i64p, _i64p, _i64p, _i64p, C.c_int,
                                         C.POINTER(C.c_void_p)]),
    "lsvd_batch_plan_execute": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "lsvd_batch_plan_destroy": (C.c_int, [_vp]),
    "lsvd_svd_batched_f64": (C.c_int, [_vp, C.c_int, _i32p, _i32p, _i64p, _vp, _i64p, _vp, _i64p, _vp, _i64p,
                                       _vp, C.c_int, _vp]),
    "lsvd_autotune_batch": (C.c_int, [_vp, C.c_int, _i32p, _i32p, C.c_int]),
    "lsvd_polar_f64": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, _vp, C.POINTER(LsvdPolarInfo)]),
    "lsvd_gemm_f64": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp]),
    "lsvd_gemm_lowp": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp]),
    "lsvd_svd_vjp_f64": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_double,
                                   _vp]),
    "lsvd_polar_vjp_f64": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp, C.c_double, _vp]),
    "lsvd_joint_params_for_n": (C.c_int, [C.c_int, C.POINTER(LsvdJointParams)]),
    "lsvd_joint_diag_sym_f64": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp, C.POINTER(LsvdJointParams), _vp,
                                          _vp, _vp]),
    "lsvd_joint_svd_f64": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, C.POINTER(LsvdJointParams),
                                     _vp, _vp, _vp, C.POINTER(LsvdJointTrace)]),
    "lsvd_adaptive_params_default": (C.c_int, [C.POINTER(LsvdAdaptiveParams)]),
    "lsvd_analyze_batched_f64": (C.c_int, [_vp, C.c_int, _i32p, _i64p, _vp, C.POINTER(LsvdAdaptiveParams),
                                           C.POINTER(LsvdTriage)]),
    "lsvd_choose_route": (C.c_int, [C.POINTER(LsvdTriage), C.POINTER(LsvdAdaptiveParams)]),
    "lsvd_polish_batched_f64": (C.c_int, [_vp, C.c_int, _i32p, _i64p, _i64p, _vp, _vp, _vp, C.c_int, _vp, _vp, _vp,
                                          _vp]),
    "lsvd_hosvd3_f64": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp,
                                  C.POINTER(LsvdTensorTrace)]),
    "lsvd_tensor_rotate3_f64": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _vp,
                                          _vp, _vp, _vp, _vp]),
    "lsvd_bss_separate_f64": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp, C.c_int, _vp, C.c_double,
                                        C.POINTER(LsvdJointParams), _vp, _vp, _vp, _vp, _vp,
                                        C.POINTER(LsvdBssTrace)]),
    "lsvd_givens_schedule_f64": (C.c_int, [_vp, C.c_int, _i32p, _i64p, _vp, C.c_int, _vp, _vp, _vp, _vp, _vp]),
    "lsvd_givens_events": (C.c_int, [C.c_int, _vp, _vp, C.POINTER(LsvdPhaseEvent)]),
    "lsvd_schedule_json": (C.c_size_t, [C.c_int, C.c_int, C.c_int, C.POINTER(LsvdPhaseEvent), C.c_int, _vp,
                                        C.c_char_p, C.c_size_t]),
    "lsvd_batch_plan_describe": (C.c_int, [_vp, _vp, C.c_int, C.POINTER(C.c_int), _vp]),
    "lsvd_batch_plan_time_buckets": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, C.c_int]),
    "lsvd_batch_plan_phase_clocks": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "lsvd_sweep_graph_chunks": (C.c_int, [C.c_int, C.c_int, C.c_int, _i32p, _i32p, C.c_int]),
    "lsvd_max_graph_nodes": (C.c_int, []),
    "lsvd_joint_layer_pairs": (C.c_int, [_vp, C.c_int, _i32p]),
    "lsvd_workspace_info": (C.c_int, [_vp, C.POINTER(C.c_size_t), C.POINTER(C.c_int)]),
    "lsvd_autotune_report": (C.c_int, [_vp, _vp]),
    "lsvd_autotune_batch_pools": (C.c_int, [_vp, C.c_int, _i32p, _i32p, _i64p, _i64p, _i64p, _i64p, C.c_int, _vp,
                                             _vp, _vp, _vp]),
    "lsvd_metrics_batched_f64": (C.c_int, [_vp, C.c_int, _i32p, _i32p, _i64p, _i64p, _i64p, _i64p, _vp, _vp, _vp,
                                            _vp, _vp, _vp]),
}


def lib():
    """Loads the HIP library; raises if it is missing (no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m lie_cliffalg_analog_svd_amd.build` "
                "(hipcc, gfx950). There is no CPU fallback.")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def status_string(status: int) -> str:
    try:
        return lib().lsvd_status_string(status).decode()
    except Exception:  # pragma: no cover - only while formatting another error
        return str(status)


def check(h, status: int) -> None:
    if status == LSVD_OK:
        return
    msg = lib().lsvd_last_error(h).decode() if h else ""
    if status == LSVD_E_SHAPE:
        raise ShapeError(status, msg)
    raise LsvdError(status, msg)
