"""Host-side initial state (``iblbm/init.py:116-265``) for the scenarios on the hot path.

Grid, relaxation time, marker ring and the initial equilibrium are computed on the host in
FP64 with the same operations in the same order as the reference, so the arrays handed to
the device carry the reference's bits.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# D2Q9 (iblbm/lbm/lattice.py:43-56)
E = np.array([[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1], [1, 1], [-1, 1], [-1, -1], [1, -1]],
             dtype=np.float64)
W = np.array([4 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 36, 1 / 36, 1 / 36, 1 / 36])
OPP = np.array([0, 3, 4, 1, 2, 7, 8, 5, 6], dtype=np.int32)


@dataclass
class HostState:
    nx: int
    ny: int
    nodenums: int
    dx: float
    dy: float
    dt: float
    tau: float
    r: float
    lattice_r: float
    in_u: float
    dens: float
    U0: np.ndarray          # initial velocity (N, 2)
    feq0: np.ndarray        # initial equilibrium (N, 9) (= initial fstar)
    Lx: np.ndarray
    Ly: np.ndarray
    Lb: int
    Larea: float
    desired_velocity: np.ndarray
    tau_minus: float | None = None
    S_diag: np.ndarray | None = None
    extra: dict = field(default_factory=dict)


def feq_host(ro: np.ndarray, U: np.ndarray, incompressible: bool = False) -> np.ndarray:
    """Equilibrium on the host, elementwise identical to ``_compute_feq_nb``
    (``iblbm/lbm/equilibrium.py:27-42``; ``x ** 2`` is lowered to ``x * x`` there)."""
    u0 = U[:, 0:1]
    u1 = U[:, 1:2]
    u2 = u0 * u0 + u1 * u1
    eU = E[None, :, 0] * u0 + E[None, :, 1] * u1
    if incompressible:
        return W[None, :] * (ro[:, None] + 3.0 * eU + 4.5 * eU * eU - 1.5 * u2)
    return ro[:, None] * W[None, :] * (1.0 + 3.0 * eU + 4.5 * eU * eU - 1.5 * u2)


def circle_markers(cx, cy, r_domain, lattice_r, dx, spacing_factor, retraction_dx):
    """Marker ring (``iblbm/init.py:210-242``, ``iblbm/physics/markers.py:60-85``)."""
    dd = 2.0 * np.pi * lattice_r / spacing_factor
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


def grid_of(cfg):
    """(nx, ny, dx, dy) of ``iblbm/init.py:133-151``."""
    if cfg.nx_formula == "standard":
        nx = int(cfg.NN * cfg.xmax - (cfg.xmax - 1))
    else:
        nx = int(cfg.NN * cfg.xmax)
    dx = 1.0 / (cfg.NN - 1)
    ny = round(cfg.ymax / dx) + 1
    dy = cfg.ymax / (ny - 1)
    if abs(dx - dy) / dx > 1e-6:
        raise ValueError(f"grid is not square: dx={dx:.8f}, dy={dy:.8f} (ymax={cfg.ymax})")
    return nx, ny, dx, dy


def initialize(cfg, perturb=None, fields: bool = True) -> HostState:
    """Initial state of ``iblbm/init.py:116-265``.

    ``perturb=(seed, amp)`` multiplies the initial velocity by (1 + amp*xi), xi ~ U(-1, 1)
    i.i.d. per node and component from ``numpy.random.default_rng(seed)``, before the
    equilibrium is formed (the seeded-perturbation inputs of the Re=100 and batched cases).
    ``fields=False`` leaves out the full-grid U0 / feq0 arrays (set to None), for instances
    whose initial state is formed on the device.
    """
    nx, ny, dx, dy = grid_of(cfg)
    N = nx * ny
    phy_char = cfg.phy_l * cfg.cylinder_D_ratio
    phy_dx = cfg.phy_l / (cfg.NN - 1)
    Cx = phy_dx / 1.0
    r = phy_char / 2.0 / cfg.phy_l
    lattice_D = phy_char / Cx
    lattice_r = lattice_D / 2.0
    lattice_nu = cfg.lattice_u * 2.0 * lattice_r / cfg.Re
    tau = 3.0 * lattice_nu + 0.5
    U = feq0 = None
    if fields:
        U = np.zeros((N, 2))
        if cfg.scenario_type == "taylor_green":  # init.py:178-191
            from .taylor_green import initial_velocity
            U = initial_velocity(cfg, nx, ny)
        elif cfg.motion_type is None:
            U[:, 0] = cfg.inflow_u
        ro = 1.0 * np.ones(N)
        if perturb is not None:
            seed, amp = perturb
            xi = np.random.default_rng(seed).uniform(-1.0, 1.0, size=U.shape)
            U = U * (1.0 + amp * xi)
        feq0 = feq_host(ro, U, cfg.incompressible_lbgk)
    cx, cy = cfg.cylinder_center
    if cx - r < 0 or cx + r > cfg.xmax or cy - r < 0 or cy + r > cfg.ymax:
        raise ValueError(f"cylinder (cx={cx}, cy={cy}, r={r}) outside [0,{cfg.xmax}]x[0,{cfg.ymax}]")
    if cfg.retraction_dx < 0.0 or cfg.retraction_dx >= lattice_r:
        raise ValueError(f"retraction_dx={cfg.retraction_dx} outside [0, lattice_r={lattice_r})")
    Lx, Ly, Lb, Larea = circle_markers(cx, cy, r, lattice_r, dx, cfg.marker_spacing_factor,
                                       cfg.retraction_dx)
    st = HostState(nx=nx, ny=ny, nodenums=N, dx=dx, dy=dy, dt=1.0, tau=tau, r=r,
                   lattice_r=lattice_r, in_u=cfg.inflow_u, dens=1.0, U0=U, feq0=feq0,
                   Lx=Lx, Ly=Ly, Lb=Lb, Larea=Larea, desired_velocity=np.zeros((Lb, 2)))
    if cfg.collision_model == "TRT":
        st.tau_minus = 0.5 + cfg.trt_magic_param / (tau - 0.5)  # trt.py:73-75
    elif cfg.collision_model == "CM_MRT":
        s_eps = cfg.mrt_s_eps if cfg.mrt_s_eps is not None else 1.4  # cm_mrt.py:171-188
        s_q = cfg.mrt_s_q if cfg.mrt_s_q is not None else 1.2
        s_nu = 1.0 / tau
        st.S_diag = np.array([0.0, 0.0, 0.0, s_nu, s_nu, s_nu, s_q, s_q, s_eps])
    elif cfg.collision_model != "BGK":
        raise KeyError(f"unknown collision_model={cfg.collision_model!r} "
                       "(registered: ['BGK', 'CM_MRT', 'TRT'])")
    return st
