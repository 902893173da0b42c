"""FP64 CPU oracle for the IB-LBM per-step hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the checker (or as
the timed CPU baseline); the shipped package never does.

It drives ``oracle/cpu_ref.c`` (a C restatement of the reference's numba kernels and NumPy
IBM, built by ``oracle/Makefile`` into ``oracle/libcpu_ref.so``) with a Python loop that
restates ``iblbm/runtime/loop.py:44-151`` for the fixed-cylinder runtime
(``iblbm/runtime/scenarios/fixed_cylinder.py``), and ``iblbm/init.py:116-265`` for the state.
Host-side quantities that the reference computes with NumPy (marker ring, Cd/Cl) are
computed here with the same NumPy calls in the same order, so they carry the same bits.

Pinned by: tests/golden/*.npz written from the reference (tests/golden/make_golden.py),
closed forms in tests/test_oracle_closed_forms.py.
"""
from __future__ import annotations

import ctypes
import math
import os
from types import SimpleNamespace

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_D = ctypes.POINTER(ctypes.c_double)
DELTA_IDS = {"peskin4pt": 0, "hat": 1}
DELTA_RADIUS = {"peskin4pt": 2, "hat": 1}


def lib():
    """Load ``libcpu_ref.so`` (built by ``make -C oracle``)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libcpu_ref.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run make -C oracle)")
        L = ctypes.CDLL(path)
        i, l, d, p = ctypes.c_int, ctypes.c_long, ctypes.c_double, _D
        sig = {
            "ref_threads": (i, [i]),
            "ref_feq": (None, [p, p, l, i, p]),
            "ref_macro": (None, [p, l, i, p, p]),
            "ref_stream": (None, [p, p, i, i, p]),
            "ref_collide_bgk": (None, [p, p, p, p, d, d, l, p]),
            "ref_collide_trt": (None, [p, p, p, p, d, d, d, l, p]),
            "ref_collide_cm": (None, [p, p, p, p, d, l, p]),
            "ref_bc_dirichlet_neumann": (None, [p, i, i, d, d]),
            "ref_bc_settling": (None, [p, i, i, d, i]),
            "ref_bc_velocity_inlet": (None, [p, p, i, i, d, d]),
            "ref_bc_open": (None, [p, p, i, i, d]),
            "ref_bc_tg": (None, [p, p, i, i]),
            "ref_delta": (d, [i, d]),
            "ref_df": (None, [p, p, p, i, p, p, p, d, d, d, d, i, i, i, i, p, p, p, p, p, p]),
            "ref_mdf": (None, [p, p, p, i, p, p, d, d, d, d, i, i, i, i, i, d, d, p, p]),
            "ref_dfc_lambda": (None, [p, p, i, d, d, d, d, i, i, i, p]),
            "ref_dfc_apply": (None, [p, p, p, i, p, p, d, d, d, i, i, i, d, p]),
            "ref_a_norm_inf": (d, [p, p, i, i, d]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(_D)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def set_threads(n: int) -> int:
    return lib().ref_threads(int(n))


# --------------------------------------------------------------------------- lattice ops

def feq(ro, U, incompressible=False):
    ro, U = _c(ro), _c(U)
    out = np.empty((ro.shape[0], 9))
    lib().ref_feq(_p(ro), _p(U), ro.shape[0], int(incompressible), _p(out))
    return out


def macroscopic(f, incompressible=False):
    f = _c(f)
    N = f.shape[0]
    ro, U = np.empty(N), np.empty((N, 2))
    lib().ref_macro(_p(f), N, int(incompressible), _p(ro), _p(U))
    return ro, U


def stream(fstar_old, f, nx, ny):
    fstar_old, f = _c(fstar_old), _c(f)
    out = np.empty_like(f)
    lib().ref_stream(_p(fstar_old), _p(f), nx, ny, _p(out))
    return out


def collide_bgk(fstar, feq_, U, fib, tau, dt=1.0):
    a = [_c(x) for x in (fstar, feq_, U, fib)]
    out = np.empty_like(a[0])
    lib().ref_collide_bgk(*map(_p, a), tau, dt, a[0].shape[0], _p(out))
    return out


def collide_trt(fstar, feq_, U, fib, s_plus, s_minus, dt=1.0):
    a = [_c(x) for x in (fstar, feq_, U, fib)]
    out = np.empty_like(a[0])
    lib().ref_collide_trt(*map(_p, a), s_plus, s_minus, dt, a[0].shape[0], _p(out))
    return out


def collide_cm(fstar, U, fib, S_diag, dt=1.0):
    a = [_c(x) for x in (fstar, U, fib, S_diag)]
    out = np.empty_like(a[0])
    lib().ref_collide_cm(*map(_p, a), dt, a[0].shape[0], _p(out))
    return out


def bc_dirichlet_neumann(fstar, nx, ny, in_u, dens=1.0):
    out = _c(fstar).copy()
    lib().ref_bc_dirichlet_neumann(_p(out), nx, ny, in_u, dens)
    return out


def bc_settling(fstar, nx, ny, dens=1.0, incompressible=False):
    out = _c(fstar).copy()
    lib().ref_bc_settling(_p(out), nx, ny, dens, int(incompressible))
    return out


def bc_velocity_inlet(fstar, U, nx, ny, in_u, dens=1.0):
    """``iblbm/boundary/velocity_inlet.py:45-157``; U is the previous step's velocity."""
    out, U = _c(fstar).copy(), _c(U)
    lib().ref_bc_velocity_inlet(_p(out), _p(U), nx, ny, in_u, dens)
    return out


def bc_open(fstar, U, nx, ny, dens=1.0):
    """``iblbm/boundary/open_boundary.py:44-162``; U is the previous step's velocity."""
    out, U = _c(fstar).copy(), _c(U)
    lib().ref_bc_open(_p(out), _p(U), nx, ny, dens)
    return out


def tg_field(X, Y, t, u0, L, nu):
    """``iblbm/diagnostics/core.py:216-239`` (same NumPy calls)."""
    decay = np.exp(-2.0 * nu * (np.pi / L) ** 2 * t)
    ux = -u0 * np.cos(np.pi * X / L) * np.sin(np.pi * Y / L) * decay
    uy = u0 * np.sin(np.pi * X / L) * np.cos(np.pi * Y / L) * decay
    return ux, uy


def tg_nu(cfg):
    """The Taylor-Green viscosity of ``tg_analytical.py:59-61`` (and ``init.py:186-187``)."""
    lattice_D = cfg.cylinder_D_ratio * (cfg.NN - 1)
    return cfg.lattice_u * lattice_D / cfg.Re


def tg_boundary_velocity(cfg, nx, ny, t):
    """The analytic velocity of every boundary node at time t, computed as
    ``tg_analytical.py:52-136`` computes it face by face: left (ny), right (ny), top (nx - 2),
    bottom (nx - 2), corners (0,0), (nx-1,0), (0,ny-1), (nx-1,ny-1); (n, 2)."""
    L_lat = 0.5 * (cfg.NN - 1)
    nu = tg_nu(cfg)
    X = np.arange(nx, dtype=float) - (nx - 1) / 2.0
    Y = np.arange(ny, dtype=float) - (ny - 1) / 2.0
    j = np.arange(ny)
    i = np.arange(1, nx - 1)
    parts = [tg_field(np.full_like(j, X[0], dtype=float), Y[j], t, cfg.tg_u0, L_lat, nu),
             tg_field(np.full_like(j, X[nx - 1], dtype=float), Y[j], t, cfg.tg_u0, L_lat, nu),
             tg_field(X[i], np.full_like(i, Y[ny - 1], dtype=float), t, cfg.tg_u0, L_lat, nu),
             tg_field(X[i], np.full_like(i, Y[0], dtype=float), t, cfg.tg_u0, L_lat, nu),
             tg_field(X[np.array([0, nx - 1, 0, nx - 1])], Y[np.array([0, 0, ny - 1, ny - 1])],
                      t, cfg.tg_u0, L_lat, nu)]
    return np.concatenate([np.stack([ux, uy], 1) for ux, uy in parts])


def bc_tg(fstar, Ub, nx, ny):
    """``iblbm/boundary/tg_analytical.py:29-136`` with the boundary velocities Ub."""
    out, Ub = _c(fstar).copy(), _c(Ub)
    lib().ref_bc_tg(_p(out), _p(Ub), nx, ny)
    return out


def tg_desired_velocity(cfg, st, t):
    """``iblbm/runtime/scenarios/taylor_green.py:21-32``: the analytic velocity at the
    markers."""
    L_lat = 0.5 * (cfg.NN - 1)
    Xc = st.Lx / st.dx - (st.nx - 1) / 2.0
    Yc = st.Ly / st.dy - (st.ny - 1) / 2.0
    ux, uy = tg_field(Xc, Yc, t, cfg.tg_u0, L_lat, tg_nu(cfg))
    ud = np.empty((len(Xc), 2))
    ud[:, 0] = ux
    ud[:, 1] = uy
    return ud


def oscillating_markers(t, Lx_c, Ly_c, lattice_u, KC, r, lattice_r):
    """``iblbm/physics/motion.py:15-52``: x(t) = A sin(2 pi f0 t); markers shifted by -x(t),
    desired velocity (x(t), 0)."""
    f0 = lattice_u / (KC * 2.0 * lattice_r)
    A = KC * 2.0 * r / (2.0 * np.pi)
    xt = A * np.sin(f0 * 2.0 * np.pi * t)
    Lx = Lx_c - xt
    Ly = Ly_c.copy()
    ud = np.empty((len(Lx_c), 2))
    ud[:, 0] = xt
    ud[:, 1] = 0.0
    return Lx, Ly, ud


def delta(delta_type, r):
    f = lib().ref_delta
    t = DELTA_IDS[delta_type]
    return np.array([f(t, float(x)) for x in np.ravel(r)]).reshape(np.shape(r))


# --------------------------------------------------------------------------- IBM ops

def direct_forcing(Lx, Ly, ud, Eux, Euy, Ro, dx, dy, dt, Larea, ny, nx, delta_type,
                   incompressible=False):
    """Restates ``iblbm/ibm/df.py:94-154``; returns (fib, Lux, Luy, Lfx, Lfy, R)."""
    Lx, Ly, ud, Eux, Euy, Ro = (_c(x) for x in (Lx, Ly, ud, Eux, Euy, Ro))
    Lb = Lx.shape[0]
    fib = np.empty((ny * nx, 2))
    o = [np.empty(Lb) for _ in range(5)]
    lib().ref_df(_p(Lx), _p(Ly), _p(ud), Lb, _p(Eux), _p(Euy), _p(Ro), dx, dy, dt, Larea,
                 ny, nx, DELTA_IDS[delta_type], int(incompressible), _p(fib), *map(_p, o))
    return (fib, *o)


def multi_direct_forcing(Lx, Ly, ud, U, ro, dx, dy, dt, Larea, ny, nx, n_iter, min_iter,
                         delta_type, omega, tol):
    """Restates ``iblbm/ibm/mdf.py:175-229``; returns (fib_total, stats dict)."""
    Lx, Ly, ud, U, ro = (_c(x) for x in (Lx, Ly, ud, U, ro))
    Lb = Lx.shape[0]
    fib = np.empty((ny * nx, 2))
    st = np.empty(4)
    lib().ref_mdf(_p(Lx), _p(Ly), _p(ud), Lb, _p(U), _p(ro), dx, dy, dt, Larea, ny, nx,
                  int(n_iter), int(min_iter), DELTA_IDS[delta_type], omega, tol, _p(fib), _p(st))
    return fib, {"iter_count": int(st[0]), "last_residual": float(st[1]),
                 "diverged": bool(st[2]), "converged": bool(st[3])}


def dfc_lambda(Lx, Ly, rho_f, Larea, dx, dy, ny, nx, delta_type):
    """Restates ``iblbm/ibm/dfc.py:126-186``."""
    Lx, Ly = _c(Lx), _c(Ly)
    lam = np.empty(Lx.shape[0])
    lib().ref_dfc_lambda(_p(Lx), _p(Ly), Lx.shape[0], rho_f, Larea, dx, dy, ny, nx,
                         DELTA_IDS[delta_type], _p(lam))
    return lam


def dfc_correction(Lx, Ly, ud, fstar, lam, dx, dy, Larea, ny, nx, delta_type, scale=1.0):
    """Restates ``iblbm/ibm/dfc.py:319-341``; returns (fstar_corrected, force)."""
    Lx, Ly, ud, lam = (_c(x) for x in (Lx, Ly, ud, lam))
    fs = _c(fstar).copy()
    force = np.empty((Lx.shape[0], 2))
    lib().ref_dfc_apply(_p(Lx), _p(Ly), _p(ud), Lx.shape[0], _p(fs), _p(lam), dx, dy, Larea,
                        ny, nx, DELTA_IDS[delta_type], scale, _p(force))
    return fs, force


def omega_from_type(Lx, Ly, delta_type, Larea, dx):
    """Restates ``iblbm/ibm/a_norm.py:204-266``: 1/||A||_inf in lattice units."""
    Lx_lu = _c(np.asarray(Lx, dtype=np.float64) / float(dx))
    Ly_lu = _c(np.asarray(Ly, dtype=np.float64) / float(dx))
    norm = lib().ref_a_norm_inf(_p(Lx_lu), _p(Ly_lu), Lx_lu.shape[0], DELTA_IDS[delta_type],
                                float(Larea))
    return 1.0 / norm


# --------------------------------------------------------------------------- state + loop

def marker_ring(cx, cy, r_domain, lattice_r, dx, spacing, retraction_dx):
    """Restates ``iblbm/init.py:211-242`` (same NumPy calls)."""
    dd = 2.0 * np.pi * lattice_r / spacing
    step = 1.0 / dd
    Ldx = np.arange(0, 1.0 + step * 0.5, step)
    Ldx = Ldx[Ldx <= 1.0 + 1e-12]
    if len(Ldx) > 1 and np.abs(Ldx[-1] - 1.0) < step * 0.5:
        Ldx = Ldx[:-1]
    theta = 2.0 * np.pi * Ldx
    r_ret = r_domain - retraction_dx * dx
    Lx = r_ret * np.cos(theta) + cx
    Ly = r_ret * np.sin(theta) + cy
    Lb = len(Lx)
    Larea = 2.0 * np.pi * (lattice_r - retraction_dx) / Lb
    return Lx, Ly, Lb, Larea


def initialize(cfg, perturb=None):
    """Restates ``iblbm/init.py:116-265`` for the fixed-cylinder scenario.

    ``perturb=(seed, amp)`` applies u <- u*(1 + amp*xi), xi ~ U(-1,1) from
    ``np.random.default_rng(seed)`` over the (N, 2) velocity, then feq = fstar = f_eq.
    """
    if getattr(cfg, "nx_formula", "standard") == "standard":
        nx = int(cfg.NN * cfg.xmax - (cfg.xmax - 1))
    else:
        nx = int(cfg.NN * cfg.xmax)
    dx = 1.0 / (cfg.NN - 1)
    ny = round(cfg.ymax / dx) + 1
    dy = cfg.ymax / (ny - 1)
    phy_l = getattr(cfg, "phy_l", 1e-3)
    phy_char = phy_l * cfg.cylinder_D_ratio
    Cx = phy_l / (cfg.NN - 1) / 1.0
    r = phy_char / 2.0 / phy_l
    lattice_r = phy_char / Cx / 2.0
    tau = 3.0 * (cfg.lattice_u * 2.0 * lattice_r / cfg.Re) + 0.5
    N = nx * ny
    U = np.zeros((N, 2))
    if getattr(cfg, "scenario_type", None) == "taylor_green":  # init.py:178-191
        L_lat = 0.5 * (cfg.NN - 1)
        XX, YY = np.meshgrid(np.arange(nx, dtype=float) - (nx - 1) / 2.0,
                             np.arange(ny, dtype=float) - (ny - 1) / 2.0)
        ux, uy = tg_field(XX.ravel(), YY.ravel(), 0.0, cfg.tg_u0, L_lat, tg_nu(cfg))
        U[:, 0] = ux
        U[:, 1] = uy
    elif getattr(cfg, "motion_type", None) is None:
        U[:, 0] = cfg.inflow_u
    ro = 1.0 * np.ones(N)
    if perturb is not None:
        seed, amp = perturb
        xi = np.random.default_rng(seed).uniform(-1.0, 1.0, size=U.shape)
        U = U * (1.0 + amp * xi)
    inc = bool(getattr(cfg, "incompressible_lbgk", False))
    fe = feq(ro, U, inc)
    cx, cy = cfg.cylinder_center
    Lx, Ly, Lb, Larea = marker_ring(cx, cy, r, lattice_r, dx, cfg.marker_spacing_factor,
                                    getattr(cfg, "retraction_dx", 0.0))
    st = SimpleNamespace(nx=nx, ny=ny, N=N, dx=dx, dy=dy, dt=1.0, tau=tau, r=r,
                         lattice_r=lattice_r, in_u=cfg.inflow_u, dens=1.0,
                         fstar=fe.copy(), feq=fe, ro=ro, U=U, fib=np.zeros((N, 2)),
                         Lx=Lx, Ly=Ly, Lb=Lb, Larea=Larea, desired_velocity=np.zeros((Lb, 2)),
                         incompressible=inc)
    model = cfg.collision_model
    if model == "TRT":
        st.tau_minus = 0.5 + cfg.trt_magic_param / (st.tau - 0.5)
    elif model == "CM_MRT":
        s_eps = cfg.mrt_s_eps if getattr(cfg, "mrt_s_eps", None) is not None else 1.4
        s_q = cfg.mrt_s_q if getattr(cfg, "mrt_s_q", None) is not None else 1.2
        s_nu = 1.0 / st.tau
        st.S_diag = np.array([0.0, 0.0, 0.0, s_nu, s_nu, s_nu, s_q, s_q, s_eps])
    return st


def collide(st, cfg):
    """Restates ``iblbm/runtime/step.py:25-41`` dispatch to the three collision kernels."""
    m = cfg.collision_model
    if m == "BGK":
        return collide_bgk(st.fstar, st.feq, st.U, st.fib, st.tau, st.dt)
    if m == "TRT":
        return collide_trt(st.fstar, st.feq, st.U, st.fib, 1.0 / st.tau, 1.0 / st.tau_minus,
                           st.dt)
    if m == "CM_MRT":
        return collide_cm(st.fstar, st.U, st.fib, st.S_diag, st.dt)
    raise KeyError(m)


def cd_cl(fib, ro, u_ref, r_lattice):
    """``iblbm/diagnostics/core.py:53-81`` (same NumPy reductions)."""
    denom = np.mean(ro) * u_ref ** 2 * r_lattice
    return float(np.abs(np.sum(fib[:, 0])) / denom), float(np.sum(fib[:, 1]) / denom)


def cd_cl_dfc(force, ro_mean, u_ref, r_lattice):
    """``iblbm/diagnostics/core.py:83-113``."""
    total = np.sum(force, axis=0)
    denom = ro_mean * u_ref ** 2 * r_lattice
    return float(np.abs(total[0])) / denom, float(total[1]) / denom


class FixedCylinder:
    """The reference time loop for a fixed cylinder (DF, MDF or DFC), one step per call.

    After ``step()`` the attributes hold what the reference's state holds at the end of that
    step: ``st.fstar`` (post-stream/BC, plus the DFC correction), ``st.ro``, ``st.U``
    (velocity-corrected for DF/MDF), ``st.fib``, ``st.feq``; ``f_post`` is the post-collision
    population the next step streams.
    """

    def __init__(self, cfg, perturb=None):
        self.cfg = cfg
        self.st = initialize(cfg, perturb)
        self.ttt = 0
        self.method = cfg.ibm_method
        self.delta = cfg.delta_type
        self.omega = None
        self.lambda_cache = None
        self.mdf_stats = None
        self.dfc_force = None
        self.Cd, self.Cl = [], []
        if self.method == "MDF":
            self.omega = omega_from_type(self.st.Lx, self.st.Ly, self.delta, self.st.Larea,
                                         self.st.dx)
        self.tg = getattr(cfg, "scenario_type", None) == "taylor_green"
        self.moving = getattr(cfg, "motion_type", None) is not None
        self.Lx_c, self.Ly_c = self.st.Lx.copy(), self.st.Ly.copy()
        self.f_post = collide(self.st, cfg)

    def step(self):
        cfg, st = self.cfg, self.st
        self.ttt += 1
        t = self.ttt * st.dt
        if self.tg and cfg.tg_with_ibm:  # taylor_green.py:48-49 (pre_step)
            st.desired_velocity = tg_desired_velocity(cfg, st, t)
        st.fstar = stream(st.fstar, self.f_post, st.nx, st.ny)
        if self.tg:  # boundary/__init__.py:14: taylor_green selects tg_analytical
            st.fstar = bc_tg(st.fstar, tg_boundary_velocity(cfg, st.nx, st.ny, t), st.nx, st.ny)
        elif cfg.bc_type == "dirichlet_neumann":
            st.fstar = bc_dirichlet_neumann(st.fstar, st.nx, st.ny, st.in_u, st.dens)
        elif cfg.bc_type == "settling_channel":
            st.fstar = bc_settling(st.fstar, st.nx, st.ny, st.dens, st.incompressible)
        elif cfg.bc_type == "velocity_inlet":
            st.fstar = bc_velocity_inlet(st.fstar, st.U, st.nx, st.ny, st.in_u, st.dens)
        elif cfg.bc_type == "open_boundary":
            st.fstar = bc_open(st.fstar, st.U, st.nx, st.ny, st.dens)
        else:
            raise KeyError(cfg.bc_type)
        st.ro, st.U = macroscopic(st.fstar, st.incompressible)
        Ro = st.ro.reshape(st.ny, st.nx)
        if self.tg and not cfg.tg_with_ibm:  # dispatch.py:38-39: no IBM at all
            st.fib = np.zeros((st.N, 2))
        elif self.method == "DF":
            st.fib = direct_forcing(st.Lx, st.Ly, st.desired_velocity,
                                    st.U[:, 0].reshape(st.ny, st.nx),
                                    st.U[:, 1].reshape(st.ny, st.nx), Ro, st.dx, st.dy, st.dt,
                                    st.Larea, st.ny, st.nx, self.delta, st.incompressible)[0]
        elif self.method == "MDF":
            st.fib, self.mdf_stats = multi_direct_forcing(
                st.Lx, st.Ly, st.desired_velocity, st.U, st.ro, st.dx, st.dy, st.dt, st.Larea,
                st.ny, st.nx, cfg.mdf_iterations, cfg.mdf_min_iterations, self.delta,
                self.omega, cfg.mdf_tolerance)
        elif self.method == "DFC":
            if self.lambda_cache is None or self.moving:  # strategies/dfc.py:32, :48-49
                self.lambda_cache = dfc_lambda(st.Lx, st.Ly, 1.0, st.Larea, st.dx, st.dy,
                                               st.ny, st.nx, self.delta)
            st.fstar, self.dfc_force = dfc_correction(
                st.Lx, st.Ly, st.desired_velocity, st.fstar, self.lambda_cache, st.dx, st.dy,
                st.Larea, st.ny, st.nx, self.delta)
            st.ro, st.U = macroscopic(st.fstar, st.incompressible)
        else:
            raise KeyError(self.method)
        if self.method in ("DF", "MDF") and not (self.tg and not cfg.tg_with_ibm):
            if st.incompressible:
                st.U = st.U + 0.5 * st.fib * st.dt
            else:
                st.U = st.U + st.fib * st.dt / (2.0 * st.ro[:, None])
        st.feq = feq(st.ro, st.U, st.incompressible)
        self.f_post = collide(st, cfg)
        if getattr(cfg, "motion_type", None) == "oscillating":  # fixed_cylinder.py:55-66
            st.Lx, st.Ly, st.desired_velocity = oscillating_markers(
                self.ttt, self.Lx_c, self.Ly_c, cfg.lattice_u, cfg.KC, st.r, st.lattice_r)
        if self.ttt % cfg.check_interval == 0:
            # fixed_cylinder.py:69-71: the oscillating cylinder is scaled by lattice_u
            u_ref = cfg.lattice_u if getattr(cfg, "motion_type", None) == "oscillating" \
                else st.in_u
            if self.tg and not cfg.tg_with_ibm:
                cd, cl = 0.0, 0.0
            elif self.method == "DFC":
                cd, cl = cd_cl_dfc(self.dfc_force, float(st.ro.mean()), u_ref, st.lattice_r)
            else:
                cd, cl = cd_cl(st.fib, st.ro, u_ref, st.lattice_r)
            self.Cd.append(cd)
            self.Cl.append(cl)

    def run(self, steps):
        for _ in range(steps):
            self.step()
        return self
