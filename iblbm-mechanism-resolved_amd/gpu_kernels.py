"""Operator-level surface with the names, argument order and return shapes of the
reference's ``iblbm/gpu_kernels.py`` wrappers (``:760-1223``), backed by ``libiblbm_hip.so``.

Arrays are NumPy (host) FP64 in the reference layouts: populations (N, 9), U (N, 2),
Eulerian fields (ny, nx), markers in domain units.  Each call uploads, runs the gfx950
kernel and downloads; the fused per-step path (``iblbm_amd.solver``) keeps state on the
device instead.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _abi
from ._abi import DELTA, dptr, f64, check

_DELTA_TYPE_IDS = dict(DELTA)
_DELTA_STENCIL_A = {"peskin4pt": 2, "hat": 1}
_BLOCK = 256


def _grid(n):
    """Workgroups of _BLOCK covering n items (the reference's launch helper,
    gpu_kernels.py:748-753); the kernels below choose their own launch shape and ignore it."""
    return ((int(n) + _BLOCK - 1) // _BLOCK,)


def _lib():
    return _abi.load()


def _delta_id(delta_type) -> int:
    if isinstance(delta_type, int):
        if delta_type not in (0, 1):
            raise KeyError(delta_type)
        return delta_type
    if delta_type not in DELTA:
        raise KeyError(delta_type)
    return DELTA[delta_type]


def feq_gpu(ro, U, incompressible=False):
    """``feq_gpu`` (gpu_kernels.py:906): (N,) , (N,2) -> (N,9)."""
    ro, U = f64(ro).reshape(-1), f64(U).reshape(-1, 2)
    out = np.empty((ro.shape[0], 9))
    check(_lib().iblbm_op_feq(dptr(ro), dptr(U), ro.shape[0], int(incompressible), dptr(out)),
          "feq_gpu")
    return out


def macroscopic_gpu(fstar, incompressible=False):
    """``macroscopic_gpu`` (gpu_kernels.py:878): (N,9) -> (ro (N,), U (N,2))."""
    fstar = f64(fstar).reshape(-1, 9)
    N = fstar.shape[0]
    ro, U = np.empty(N), np.empty((N, 2))
    check(_lib().iblbm_op_macroscopic(dptr(fstar), N, int(incompressible), dptr(ro), dptr(U)),
          "macroscopic_gpu")
    return ro, U


def streaming_gpu(fstar_old, f, nx, ny):
    """``streaming_gpu`` (gpu_kernels.py:846) with the push semantics of streaming.py:38-77."""
    fstar_old, f = f64(fstar_old).reshape(-1, 9), f64(f).reshape(-1, 9)
    out = np.empty_like(f)
    check(_lib().iblbm_op_streaming(dptr(fstar_old), dptr(f), int(nx), int(ny), dptr(out)),
          "streaming_gpu")
    return out


def _collision(model, fstar, feq, U, fib, tau, tau_minus, S_diag, dt):
    fstar = f64(fstar).reshape(-1, 9)
    N = fstar.shape[0]
    feq = f64(feq).reshape(-1, 9) if feq is not None else None
    U, fib = f64(U).reshape(-1, 2), f64(fib).reshape(-1, 2)
    S = f64(S_diag).reshape(9) if S_diag is not None else None
    out = np.empty_like(fstar)
    check(_lib().iblbm_op_collision(model, dptr(fstar), dptr(feq) if feq is not None else None,
                                    dptr(U), dptr(fib), N, float(tau), float(tau_minus),
                                    dptr(S) if S is not None else None, float(dt), dptr(out)),
          "collision_gpu")
    return out


def collision_gpu(fstar, feq, U, fib, tau, dt):
    """BGK + Guo (gpu_kernels.py:760; arithmetic of bgk.py:25-40)."""
    return _collision(0, fstar, feq, U, fib, tau, 1.0, None, dt)


def collision_trt_gpu(fstar, feq, U, fib, tau, tau_minus, dt):
    """TRT + Guo (gpu_kernels.py:787; arithmetic of trt.py:38-67)."""
    return _collision(1, fstar, feq, U, fib, tau, tau_minus, None, dt)


def collision_cm_mrt_gpu(fstar, feq, U, fib, S_diag, dt):
    """CM-MRT + Guo (gpu_kernels.py:815; arithmetic of cm_mrt.py:52-165)."""
    return _collision(2, fstar, None, U, fib, 1.0, 1.0, S_diag, dt)


def boundary_gpu(bc_type, fstar, nx, ny, in_u, dens=1.0, incompressible=False):
    """apply_bc_dirichlet_neumann / apply_bc_settling_channel on a copy of fstar."""
    if bc_type not in _abi.BC:
        raise KeyError(bc_type)
    out = f64(fstar).reshape(-1, 9).copy()
    check(_lib().iblbm_op_boundary(_abi.BC[bc_type], dptr(out), int(nx), int(ny), float(in_u),
                                   float(dens), int(incompressible)), "boundary_gpu")
    return out


def ibm_direct_forcing_gpu(Lx, Ly, desired_vel, Eux, Euy, Ro, dx, dy, dt, Larea, ny, nx,
                           delta_type="peskin4pt", return_interp_vel=False,
                           incompressible=False):
    """``ibm_direct_forcing_gpu`` (gpu_kernels.py:956): fib (N,2) [, Lux, Luy]."""
    Lx, Ly = f64(Lx).reshape(-1), f64(Ly).reshape(-1)
    ud = f64(desired_vel).reshape(-1, 2)
    Eux, Euy, Ro = (f64(a).reshape(-1) for a in (Eux, Euy, Ro))
    Lb = Lx.shape[0]
    fib = np.empty((int(ny) * int(nx), 2))
    Lux, Luy = np.empty(Lb), np.empty(Lb)
    check(_lib().iblbm_op_direct_forcing(dptr(Lx), dptr(Ly), dptr(ud), Lb, dptr(Eux), dptr(Euy),
                                         dptr(Ro), float(dx), float(dy), float(dt), float(Larea),
                                         int(ny), int(nx), _delta_id(delta_type),
                                         int(incompressible), dptr(fib), dptr(Lux), dptr(Luy)),
          "ibm_direct_forcing_gpu")
    return (fib, Lux, Luy) if return_interp_vel else fib


def ibm_direct_forcing_fields_gpu(Lx, Ly, desired_vel, Eux, Euy, Ro, dx, dy, dt, Larea, ny,
                                  nx, delta_type="peskin4pt", incompressible=False):
    """``ibm_direct_forcing_fields_gpu`` (gpu_kernels.py:1045): (Efx, Efy, Lux, Luy)."""
    fib, Lux, Luy = ibm_direct_forcing_gpu(Lx, Ly, desired_vel, Eux, Euy, Ro, dx, dy, dt, Larea,
                                           ny, nx, delta_type, True, incompressible)
    return (fib[:, 0].reshape(int(ny), int(nx)).copy(), fib[:, 1].reshape(int(ny), int(nx)).copy(),
            Lux, Luy)


def velocity_correction_gpu(U, fib, ro, dt, incompressible=False):
    """``velocity_correction_gpu`` (gpu_kernels.py:930), in place on U (N, 2): u + fib dt / (2
    rho), or u + 0.5 fib dt when incompressible (the arithmetic of dispatch.py:30-33)."""
    if not isinstance(U, np.ndarray) or U.dtype != np.float64:
        raise ValueError("velocity_correction_gpu updates U in place: U must be a float64 array")
    Uc = U if U.flags.c_contiguous else np.ascontiguousarray(U)
    Uc = Uc.reshape(-1, 2)
    fib, ro = f64(fib).reshape(-1, 2), f64(ro).reshape(-1)
    check(_lib().iblbm_op_velocity_correction(dptr(Uc), dptr(fib), dptr(ro), Uc.shape[0],
                                              float(dt), int(bool(incompressible))),
          "velocity_correction_gpu")
    if Uc is not U and not np.shares_memory(Uc, U):
        U[...] = Uc.reshape(U.shape)


def ibm_multi_direct_forcing_gpu(Lx, Ly, desired_vel, U, ro, dx, dy, dt, Larea, ny, nx,
                                 n_iter=10, min_iter=5, delta_type="peskin4pt", omega=1.0,
                                 tol=1e-4):
    """The MDF loop of ``iblbm/ibm/mdf.py:61-229`` (no rotation coupling) in one workgroup:
    returns (fib_total (N,2), stats dict)."""
    Lx, Ly = f64(Lx).reshape(-1), f64(Ly).reshape(-1)
    ud, U, ro = f64(desired_vel).reshape(-1, 2), f64(U).reshape(-1, 2), f64(ro).reshape(-1)
    Lb = Lx.shape[0]
    fib = np.empty((int(ny) * int(nx), 2))
    st = _abi.MdfStats()
    check(_lib().iblbm_op_multi_direct_forcing(
        dptr(Lx), dptr(Ly), dptr(ud), Lb, dptr(U), dptr(ro), float(dx), float(dy), float(dt),
        float(Larea), int(ny), int(nx), int(n_iter), int(min_iter), _delta_id(delta_type),
        float(omega), float(tol), dptr(fib), ctypes.byref(st)), "ibm_multi_direct_forcing_gpu")
    return fib, {"iter_count": st.iter_count, "last_residual": st.last_residual,
                 "diverged": bool(st.diverged), "converged": bool(st.converged)}


def compute_lambda_gpu(Lx, Ly, rho_f, Larea, dx, dy, ny, nx, delta_type):
    """``compute_lambda_gpu`` (gpu_kernels.py:1091): lambda (Lb,)."""
    Lx, Ly = f64(Lx).reshape(-1), f64(Ly).reshape(-1)
    lam = np.empty(Lx.shape[0])
    check(_lib().iblbm_op_dfc_lambda(dptr(Lx), dptr(Ly), Lx.shape[0], float(rho_f), float(Larea),
                                     float(dx), float(dy), int(ny), int(nx),
                                     _delta_id(delta_type), dptr(lam)), "compute_lambda_gpu")
    return lam


def dfc_correction_gpu(Lx, Ly, desired_vel, fstar, lambda_k, dx, dy, Larea, ny, nx, delta_type,
                       lattice):
    """``dfc_correction_gpu`` (gpu_kernels.py:1147-1223): (Ef_corr (N, 9), dfc_force (Lb, 2)).

    Ef_corr is the Eulerian correction alone (zero outside the marker window), which the
    caller adds to fstar (``iblbm/ibm/dfc.py:307-316``); ``lattice`` is accepted for the
    signature (the D2Q9 constants are compiled in).  lambda_k None recomputes it."""
    Lx, Ly = f64(Lx).reshape(-1), f64(Ly).reshape(-1)
    ud = f64(desired_vel).reshape(-1, 2)
    fs = f64(fstar).reshape(-1, 9)
    lam = f64(lambda_k).reshape(-1) if lambda_k is not None else None
    ef = np.empty_like(fs)
    force = np.empty((Lx.shape[0], 2))
    check(_lib().iblbm_op_dfc_correction_field(
        dptr(Lx), dptr(Ly), dptr(ud), Lx.shape[0], dptr(fs),
        dptr(lam) if lam is not None else None, float(dx), float(dy), float(Larea), int(ny),
        int(nx), _delta_id(delta_type), dptr(ef), dptr(force)), "dfc_correction_gpu")
    return ef, force


def apply_dfc_correction_gpu(Lx, Ly, desired_vel, fstar, dx, dy, Larea, ny, nx, delta_type,
                             lattice=None, lambda_cache=None, correction_scale=1.0):
    """``apply_dfc_correction`` (``iblbm/ibm/dfc.py:256-342``) in one device call, with the CPU
    path's arithmetic (the scale applied to lambda * deviation before spreading):
    (fstar_corrected (N, 9), dfc_force (Lb, 2), lambda_k (Lb,))."""
    Lx, Ly = f64(Lx).reshape(-1), f64(Ly).reshape(-1)
    ud = f64(desired_vel).reshape(-1, 2)
    fs = f64(fstar).reshape(-1, 9).copy()
    lam = (f64(lambda_cache).reshape(-1) if lambda_cache is not None else
           compute_lambda_gpu(Lx, Ly, 1.0, Larea, dx, dy, ny, nx, delta_type))
    force = np.empty((Lx.shape[0], 2))
    check(_lib().iblbm_op_dfc_correction(dptr(Lx), dptr(Ly), dptr(ud), Lx.shape[0], dptr(fs),
                                         dptr(lam), float(dx), float(dy), float(Larea), int(ny),
                                         int(nx), _delta_id(delta_type), float(correction_scale),
                                         dptr(force)), "apply_dfc_correction_gpu")
    return fs, force, lam


class _MarkerKernel:
    """A reference-private kernel symbol (``_ibm_interp_force_kern``, ``_ibm_spread_kern``,
    gpu_kernels.py:739-740) as ``ibm/forcing.py:57-63`` and ``:157-164`` call it:
    ``kern(grid, block, args)`` with the reference's argument tuple, outputs written into the
    arrays the caller passes.  grid and block are accepted and ignored."""

    def __init__(self, name, fn):
        self.name, self._fn = name, fn

    def __call__(self, grid, block, args):
        self._fn(*args)

    def __repr__(self):
        return f"<marker kernel {self.name}>"


def _out(a, n):
    if not isinstance(a, np.ndarray) or a.dtype != np.float64 or not a.flags.c_contiguous \
            or a.size != n:
        raise ValueError("output arrays must be C-contiguous float64 arrays of the right size")
    return a


def _interp_force(Lx, Ly, desired_vel, Eux, Euy, Ro, dx, dy, dt, nx, ny, Lb, delta_type_id,
                  stencil_a, incompressible, Lfx, Lfy, Lux, Luy):
    """ibm_interp_force_kernel (gpu_kernels.py:236-286): per marker the interpolated velocity
    and the force 2 rho~ (ud - U~) / dt (incompressible: 2 (ud - U~) / dt)."""
    Lb = int(Lb)
    Lx, Ly = f64(Lx).reshape(-1), f64(Ly).reshape(-1)
    ud = f64(desired_vel).reshape(-1, 2)
    Eux, Euy, Ro = (f64(a).reshape(-1) for a in (Eux, Euy, Ro))
    if _DELTA_STENCIL_A[{0: "peskin4pt", 1: "hat"}[int(delta_type_id)]] != int(stencil_a):
        raise ValueError("stencil_a does not match delta_type_id")
    outs = [_out(a, Lb) for a in (Lfx, Lfy, Lux, Luy)]
    check(_lib().iblbm_op_interp_force(dptr(Lx), dptr(Ly), dptr(ud), Lb, dptr(Eux), dptr(Euy),
                                       dptr(Ro), float(dx), float(dy), float(dt), int(ny),
                                       int(nx), _delta_id(int(delta_type_id)),
                                       int(bool(incompressible)), *map(dptr, outs)),
          "_ibm_interp_force_kern")


def _spread(Lfx, Lfy, Lx, Ly, dx, dy, Larea, nx, ny, Lb, delta_type_id, stencil_a, Efx, Efy):
    """ibm_spread_kernel (gpu_kernels.py:294-330): adds sum_k Lf_k W ds to Efx, Efy (ny, nx)
    in place, in np.add.at order (exact for the zeroed fields the reference passes)."""
    Lb = int(Lb)
    Lfx, Lfy, Lx, Ly = (f64(a).reshape(-1) for a in (Lfx, Lfy, Lx, Ly))
    if _DELTA_STENCIL_A[{0: "peskin4pt", 1: "hat"}[int(delta_type_id)]] != int(stencil_a):
        raise ValueError("stencil_a does not match delta_type_id")
    Efx, Efy = _out(Efx, int(nx) * int(ny)), _out(Efy, int(nx) * int(ny))
    check(_lib().iblbm_op_spread(dptr(Lfx), dptr(Lfy), dptr(Lx), dptr(Ly), Lb, float(dx),
                                 float(dy), float(Larea), int(ny), int(nx),
                                 _delta_id(int(delta_type_id)), dptr(Efx), dptr(Efy)),
          "_ibm_spread_kern")


_ibm_interp_force_kern = _MarkerKernel("_ibm_interp_force_kern", _interp_force)
_ibm_spread_kern = _MarkerKernel("_ibm_spread_kern", _spread)


def estimate_optimal_omega_from_type(Lx, Ly, delta_type, Larea, dx):
    """``iblbm/ibm/a_norm.py:204-266``: 1/||A||_inf (host C++ in the library)."""
    Lx, Ly = f64(Lx).reshape(-1), f64(Ly).reshape(-1)
    if delta_type not in DELTA:
        raise KeyError(f"a_norm: unknown delta_type '{delta_type}'")
    if float(dx) <= 0.0:
        raise ValueError(f"a_norm: non-positive dx={dx}")
    return float(_lib().iblbm_omega_estimate(dptr(Lx), dptr(Ly), Lx.shape[0], DELTA[delta_type],
                                             float(Larea), float(dx)))
