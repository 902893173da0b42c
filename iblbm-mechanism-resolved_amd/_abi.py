"""ctypes binding of ``libiblbm_hip.so`` (declarations in ``include/iblbm_hip.h``).

Loading fails loudly: there is no CPU fallback behind this module.  The library is built
in-tree by ``iblbm-mechanism-resolved_amd/build.py`` (``__graft_entry__.build()``).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libiblbm_hip.so")

EINVAL, EKEY, EDEVICE, ENOMEM = -1, -2, -3, -4

COLL = {"BGK": 0, "TRT": 1, "CM_MRT": 2}
BC = {"dirichlet_neumann": 0, "settling_channel": 1, "velocity_inlet": 2, "open_boundary": 3,
      "tg_analytical": 4}
IBM = {None: 0, "NONE": 0, "DF": 1, "MDF": 2, "DFC": 3}
DELTA = {"peskin4pt": 0, "hat": 1}
DELTA_RADIUS = {"peskin4pt": 2, "hat": 1}
PRED = {"off": 0, "collect": 1, "warm": 2, "learned": 3}
MODES = {"auto": -1, "fused": 0, "two_streams": 1, "serial": 2}
MLP_IN, MLP_OUT, MLP_NP, MLP_NORM = 8, 2, 4866, 18
FP64, FP32_DEV = 0, 1


class Case(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("collision", ctypes.c_int32), ("bc", ctypes.c_int32),
        ("ibm", ctypes.c_int32), ("delta", ctypes.c_int32),
        ("incompressible", ctypes.c_int32), ("precision", ctypes.c_int32),
        ("tau", ctypes.c_double), ("tau_minus", ctypes.c_double),
        ("s_diag", ctypes.c_double * 9),
        ("dt", ctypes.c_double), ("dx", ctypes.c_double), ("dy", ctypes.c_double),
        ("in_u", ctypes.c_double), ("dens", ctypes.c_double),
        ("mdf_iterations", ctypes.c_int32), ("mdf_min_iterations", ctypes.c_int32),
        ("mdf_tolerance", ctypes.c_double), ("mdf_omega", ctypes.c_double),
        ("lb_capacity", ctypes.c_int32), ("box_capacity", ctypes.c_int32),
        ("fixed_body", ctypes.c_int32), ("reserved", ctypes.c_int32),
    ]


class Schedule(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int32), ("bulk_threads", ctypes.c_int32),
                ("bulk_rows", ctypes.c_int32), ("bulk_waves", ctypes.c_int32),
                ("step_rows", ctypes.c_int32), ("step_order", ctypes.c_int32),
                ("row_layout", ctypes.c_int32), ("store_mode", ctypes.c_int32),
                ("bulk_delay", ctypes.c_int32)]


class MdfStats(ctypes.Structure):
    _fields_ = [("iter_count", ctypes.c_int32), ("diverged", ctypes.c_int32),
                ("converged", ctypes.c_int32), ("guard_trips", ctypes.c_int32),
                ("last_residual", ctypes.c_double), ("omega_used", ctypes.c_double)]


class Diag(ctypes.Structure):
    _fields_ = [("rho_sum", ctypes.c_double), ("force", ctypes.c_double * 2),
                ("torque_fib", ctypes.c_double), ("torque_q", ctypes.c_double),
                ("step", ctypes.c_int64)]


_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_VP = ctypes.c_void_p
_F = ctypes.POINTER(ctypes.c_float)
_i, _d, _l = ctypes.c_int, ctypes.c_double, ctypes.c_int64

# name -> (restype, argtypes)
SIGNATURES = {
    "iblbm_ctx_create": (_i, [_i, ctypes.POINTER(_VP)]),
    "iblbm_ctx_destroy": (_i, [_VP]),
    "iblbm_last_error": (ctypes.c_char_p, []),
    "iblbm_device_name": (_i, [_i, ctypes.c_char_p, _i]),
    "iblbm_instance_add": (_i, [_VP, ctypes.POINTER(Case), _I]),
    "iblbm_upload_f": (_i, [_VP, _i, _D]),
    "iblbm_download_f": (_i, [_VP, _i, _D]),
    "iblbm_set_markers": (_i, [_VP, _i, _D, _D, _D, _i, _d]),
    "iblbm_set_torque_center": (_i, [_VP, _i, _d, _d]),
    "iblbm_set_particles": (_i, [_VP, _i, _i, _I, _D]),
    "iblbm_download_particle_fib": (_i, [_VP, _i, _i, _D, _D, _I]),
    "iblbm_download_marker_force": (_i, [_VP, _i, _D]),
    "iblbm_set_markers_batch": (_i, [_VP, _i, _I, _I, _D, _D, _D, _D, _D]),
    "iblbm_download_marker_forces_batch": (_i, [_VP, _i, _I, _D, _l]),
    "iblbm_marker_box": (_i, [_D, _D, _i, _d, _d, _i, _i, _i, _I]),
    "iblbm_set_boundary_velocity": (_i, [_VP, _i, _D]),
    "iblbm_set_analytic_boundary": (_i, [_VP, _i, _D, _i, _D, _i]),
    "iblbm_set_desired_velocity": (_i, [_VP, _i, _D]),
    "iblbm_step": (_i, [_VP, _i, _i]),
    "iblbm_sync": (_i, [_VP]),
    "iblbm_graph_run": (_i, [_VP, _i, _i]),
    "iblbm_step_persistent": (_i, [_VP, _i, _i]),
    "iblbm_set_concurrent": (_i, [_VP, _i]),
    "iblbm_step_mode": (_i, [_VP, _I]),
    "iblbm_get_schedule": (_i, [_VP, ctypes.POINTER(Schedule)]),
    "iblbm_set_schedule": (_i, [_VP, ctypes.POINTER(Schedule)]),
    "iblbm_row_layout_in_use": (_i, [_VP, _I]),
    "iblbm_autotune": (_i, [_VP, _i, _D, _I]),
    "iblbm_init_flow": (_i, [_VP, _i, _d, _d, _d, _d, ctypes.c_uint64]),
    "iblbm_get_mdf_stats": (_i, [_VP, _i, ctypes.POINTER(MdfStats)]),
    "iblbm_get_diag": (_i, [_VP, _i, ctypes.POINTER(Diag)]),
    "iblbm_diag_fetch": (_i, [_VP, _i]),
    "iblbm_get_diag_slot": (_i, [_VP, _i, _i, ctypes.POINTER(Diag)]),
    "iblbm_event_record": (_i, [_VP, _i]),
    "iblbm_event_elapsed": (_i, [_VP, _i, _i, _D]),
    "iblbm_download_fib": (_i, [_VP, _i, _D]),
    "iblbm_download_state": (_i, [_VP, _i, _D, _D, _D]),
    "iblbm_download_state_box": (_i, [_VP, _i, _i, _i, _i, _i, _D, _D]),
    "iblbm_download_dfc_force": (_i, [_VP, _i, _D]),
    "iblbm_download_window": (_i, [_VP, _i, _D, _D, _I]),
    "iblbm_step_counter": (_l, [_VP]),
    "iblbm_last_timing": (_i, [_VP, _D, _I]),
    "iblbm_omega_estimate": (_d, [_D, _D, _i, _i, _d, _d]),
    "iblbm_op_feq": (_i, [_D, _D, _l, _i, _D]),
    "iblbm_op_macroscopic": (_i, [_D, _l, _i, _D, _D]),
    "iblbm_op_streaming": (_i, [_D, _D, _i, _i, _D]),
    "iblbm_op_collision": (_i, [_i, _D, _D, _D, _D, _l, _d, _d, _D, _d, _D]),
    "iblbm_op_boundary": (_i, [_i, _D, _i, _i, _d, _d, _i]),
    "iblbm_op_direct_forcing": (_i, [_D, _D, _D, _i, _D, _D, _D, _d, _d, _d, _d, _i, _i, _i, _i,
                                     _D, _D, _D]),
    "iblbm_op_multi_direct_forcing": (_i, [_D, _D, _D, _i, _D, _D, _d, _d, _d, _d, _i, _i, _i,
                                           _i, _i, _d, _d, _D, ctypes.POINTER(MdfStats)]),
    "iblbm_op_multi_direct_forcing_noalloc": (_i, [_VP, _VP, _VP, _i, _VP, _VP, _d, _d, _d,
                                                   _d, _i, _i, _i, _i, _i, _d, _d, _VP,
                                                   ctypes.POINTER(MdfStats), _VP]),
    "iblbm_op_dfc_lambda": (_i, [_D, _D, _i, _d, _d, _d, _d, _i, _i, _i, _D]),
    "iblbm_op_dfc_correction_field": (_i, [_D, _D, _D, _i, _D, _D, _d, _d, _d, _i, _i, _i,
                                           _D, _D]),
    "iblbm_op_velocity_correction": (_i, [_D, _D, _D, _l, _d, _i]),
    "iblbm_op_interp_force": (_i, [_D, _D, _D, _i, _D, _D, _D, _d, _d, _d, _i, _i, _i, _i,
                                   _D, _D, _D, _D]),
    "iblbm_op_spread": (_i, [_D, _D, _D, _D, _i, _d, _d, _d, _i, _i, _i, _D, _D]),
    "iblbm_op_dfc_correction": (_i, [_D, _D, _D, _i, _D, _D, _d, _d, _d, _i, _i, _i, _d, _D]),
    "iblbm_predictor_mode": (_i, [_VP, _i, _i, _l]),
    "iblbm_predictor_samples": (_i, [_VP, _i, ctypes.POINTER(ctypes.c_int64)]),
    "iblbm_predictor_samples_written": (_i, [_VP, _i, ctypes.POINTER(ctypes.c_int64)]),
    "iblbm_predictor_data": (_i, [_VP, _i, _l, _l, _F]),
    "iblbm_predictor_train": (_i, [_VP, _i, _i, _d, ctypes.c_uint32, _F]),
    "iblbm_predictor_params": (_i, [_VP, _i, _F, _F]),
    "iblbm_predictor_set_params": (_i, [_VP, _i, _F, _F]),
    "iblbm_predictor_stats": (_i, [_VP, _i, _D, _i]),
    "iblbm_predictor_previous_force": (_i, [_VP, _i, _D, _i]),
    "iblbm_mlp_loss_grad": (_i, [_F, _F, _F, _i, _F, _F]),
    "iblbm_mlp_infer_bf16": (_i, [_F, _F, _i, _F]),
}

_LIB = None


def load(path: str = LIB_PATH):
    """Load the HIP library (raises RuntimeError if it is missing: no silent fallback)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise RuntimeError(
            f"libiblbm_hip.so not built ({path}); run __graft_entry__.build() first")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def check(rc: int, what: str = "") -> None:
    """Map a negative status to the exception class the reference raises."""
    if rc == 0:
        return
    msg = (_LIB.iblbm_last_error() or b"").decode(errors="replace") if _LIB else ""
    text = f"{what}: {msg}" if what else msg
    if rc == EKEY:
        raise KeyError(text)
    if rc == EINVAL:
        raise ValueError(text)
    if rc == ENOMEM:
        raise MemoryError(text)
    raise RuntimeError(text)


def dptr(a: np.ndarray):
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise ValueError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(_D)


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def fptr(a: np.ndarray):
    if a.dtype != np.float32 or not a.flags.c_contiguous:
        raise ValueError("expected a C-contiguous float32 array")
    return a.ctypes.data_as(_F)
