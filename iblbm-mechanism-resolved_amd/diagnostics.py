"""Result quantities of a run: Strouhal number and boundary slip error.

``compute_strouhal`` restates ``iblbm/diagnostics/core.py:138-214`` (host NumPy on the Cl
history).  ``compute_slip_error`` restates ``iblbm/diagnostics/ibm_quality.py:53-187``: the
diagnostic marker ring is built on the host and the velocity is interpolated at the markers
by the gfx950 interpolation kernel (the same stencil loop and summation order as the
reference's ``_interpolate_to_markers``).  Cd/Cl come from the fused step's FP64 reductions
(``solver.FixedCylinder.cd_cl``); the settling-particle records from
``sedimentation.record_state``.
"""
from __future__ import annotations

import numpy as np

from . import gpu_kernels


def compute_strouhal(Cl_history, steps, D_lattice: float, u_ref: float,
                     check_interval: int) -> float:
    """St = f_shed D / U_ref from the second half of the Cl series (core.py:138-214).

    The FFT peak (DC removed) of the mean-free second half gives f_peak.  A series that is
    never below -1e-10 is treated as |Cl| (rectified), whose peak sits at 2 f_shed or at its
    alias: the direct reading f_peak / 2 is kept when it gives St >= 0.1, otherwise the
    de-aliased (f_sampling - f_peak) / 2.
    """
    Cl_arr = np.asarray(Cl_history, dtype=float)
    N = len(Cl_arr)
    half = Cl_arr[N // 2:]
    is_rectified = bool(np.all(Cl_arr >= -1e-10))
    half_detrended = half - np.mean(half)
    power = np.abs(np.fft.rfft(half_detrended)) ** 2
    power[0] = 0
    peak_idx = int(np.argmax(power))
    f_peak = peak_idx / (len(half) * check_interval)
    if is_rectified:
        f_sampling = 1.0 / check_interval
        st_direct = (f_peak / 2) * D_lattice / u_ref
        if st_direct >= 0.1:
            f_shed = f_peak / 2
        else:
            f_shed = (f_sampling - f_peak) / 2
    else:
        f_shed = f_peak
    return f_shed * D_lattice / u_ref


def cylinder_markers(cx: float, cy: float, r: float, NN: int,
                     marker_spacing_factor: float = 0.5):
    """Diagnostic ring on the surface itself, no retraction (ibm_quality.py:53-87)."""
    lattice_r = r * (NN - 1)
    dd = 2.0 * np.pi * lattice_r / marker_spacing_factor
    step = 1.0 / dd
    Ldx = np.arange(0, 1.0 + step * 0.5, step)
    Ldx = Ldx[Ldx <= 1.0 + 1e-12]
    if len(Ldx) > 1 and np.abs(Ldx[-1] - 1.0) < step * 0.5:
        Ldx = Ldx[:-1]
    theta = 2.0 * np.pi * Ldx
    return r * np.cos(theta) + cx, r * np.sin(theta) + cy, theta


def interpolate_to_markers(Eux, Euy, Lx, Ly, dx: float, dy: float, delta_type: str):
    """U~(X_m) = sum_x u(x) phi phi on the device (ibm_quality.py:89-130)."""
    Eux, Euy = np.asarray(Eux, dtype=float), np.asarray(Euy, dtype=float)
    ny, nx = Eux.shape
    Lb = len(Lx)
    _, Lux, Luy = gpu_kernels.ibm_direct_forcing_gpu(
        np.asarray(Lx, dtype=float), np.asarray(Ly, dtype=float), np.zeros((Lb, 2)), Eux, Euy,
        np.ones((ny, nx)), dx, dy, 1.0, 1.0, ny, nx, delta_type=delta_type,
        return_interp_vel=True)
    return Lux, Luy


def compute_slip_error(Eux, Euy, dx: float, dy: float, cx: float, cy: float, D: float,
                       NN: int, delta_type: str, u_target_x=None, u_target_y=None,
                       u_ref: float = 0.1) -> dict:
    """Mean and max boundary slip |U~ - U_target| / U_ref (ibm_quality.py:132-187)."""
    r = D / 2.0
    Lx, Ly, theta = cylinder_markers(cx, cy, r, NN)
    Lb = len(Lx)
    Lux, Luy = interpolate_to_markers(Eux, Euy, Lx, Ly, dx, dy, delta_type)
    if u_target_x is None:
        u_target_x = np.zeros(Lb)
    if u_target_y is None:
        u_target_y = np.zeros(Lb)
    slip_x = Lux - u_target_x
    slip_y = Luy - u_target_y
    slip_mag = np.sqrt(slip_x ** 2 + slip_y ** 2)
    return {
        "mean": float(np.mean(slip_mag) / u_ref),
        "max": float(np.max(slip_mag) / u_ref),
        "per_marker": slip_mag / u_ref,
        "theta": theta,
    }
