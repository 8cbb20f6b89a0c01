"""MI355X-native (gfx950) dense SVD path of skovnats/lie_cliffalg_analog_svd.

Hand-written HIP kernels behind a C ABI (``include/lie_svd_amd.h``,
``liblsvd_amd.so``), with a Python mirror of the reference's Rust API:
``LieSvdSmall``, ``LieSvdMicro``, ``LieSvd``, ``eigh_jacobi_full``, the
batch plans, ``LieSvdJoint`` and the rotor-stream export
(``HardwareSchedule``), tensor HO-SVD (``LieSvdTensor``), Phase-BSS (``LieSvdBss``).  ``profiles`` and ``metrics`` restate the reference's input
generators and residual metrics (src/lib.rs:133-388).
"""
from ._lib import (LSVD_PREC_BF16, LSVD_PREC_F32, LSVD_PREC_F64, LSVD_PREC_F64_ONESIDED, LSVD_PREC_F64_SCALED,
                   LSVD_PTR_DEVICE,
                   LSVD_PTR_HOST, LsvdError, ShapeError)
from .handle import BatchPlan, Handle, default_handle, pack_offsets
from .solvers import (MICRO_MAX_N, SMALL_MAX_N, LieSvd, LieSvdMicro, LieSvdSmall, digital_polish,
                      digital_polish_batched, eigh_jacobi_full, run_small, svd_batched)
from . import metrics, profiles
from .polar import newton_schulz_polar
from .adjoint import polar_vjp, svd_loss_grad, svd_vjp
from .joint import (JointDiagonalizationParams, JointDiagonalizationTrace, JointSvdTrace, LieSvdJoint,
                    diagonalize_families, svd_families)
from .adaptive import (AdaptiveRoute, AdaptiveTrace, AdaptiveTriage, LieSvdAdaptive, LieSvdAdaptiveParams,
                       analyze_batch, analyze_matrix, choose_route)
from .tensor import (LieSvdTensor, TensorPhaseSvd3, TensorPhaseTrace, hosvd3_batched, reconstruct_hosvd3,
                     rotate_tensor3_batched, tensor_relative_error)
from .bss import (LieSvdBss, PhaseBssParams, PhaseBssResult, PhaseBssTrace, channel_phase_coherence,
                  estimate_sir_db, separate_batched)
from .schedule import HardwarePhaseEvent, HardwareSchedule, HardwareTarget, givens_batch, schedules_from_orthogonal

__version__ = "0.1.0"
