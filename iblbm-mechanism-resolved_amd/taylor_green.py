"""Taylor-Green vortex scenario: the analytic field and the tables the device boundary uses.

The reference (``iblbm/runtime/scenarios/taylor_green.py``, ``iblbm/boundary/tg_analytical.py``,
``scenarios/taylor_green.py``) runs a decaying vortex in the square [-L, L]^2 with every face
held at the analytic velocity

    u_x = -u0 cos(pi x / L) sin(pi y / L) d(t),   u_y = u0 sin(pi x / L) cos(pi y / L) d(t),
    d(t) = exp(-2 nu (pi / L)^2 t)                 (iblbm/diagnostics/core.py:216-239),

optionally with a cylinder whose markers follow the analytic velocity.  The host forms the
time-independent products (-u0 cos .. sin .. and u0 sin .. cos ..) and the decay d(t) with the
reference's NumPy expressions, face by face as ``tg_analytical.py:52-136`` does; the device
multiplies them (one IEEE product, the reference's last operation), so its boundary velocities
carry the reference's bits.
"""
from __future__ import annotations

import numpy as np


def tg_nu(cfg) -> float:
    """``tg_analytical.py:59-61``, ``init.py:186-187``."""
    lattice_D = cfg.cylinder_D_ratio * (cfg.NN - 1)
    return cfg.lattice_u * lattice_D / cfg.Re


def tg_half_width(cfg) -> float:
    return 0.5 * (cfg.NN - 1)


def analytic_field(X, Y, t, u0, L, nu):
    """``tg_analytical_velocity_field`` (``iblbm/diagnostics/core.py:216-239``)."""
    decay = np.exp(-2.0 * nu * (np.pi / L) ** 2 * t)
    ux = -u0 * np.cos(np.pi * X / L) * np.sin(np.pi * Y / L) * decay
    uy = u0 * np.sin(np.pi * X / L) * np.cos(np.pi * Y / L) * decay
    return ux, uy


def _coef(X, Y, u0, L):
    return (-u0 * np.cos(np.pi * X / L) * np.sin(np.pi * Y / L),
            u0 * np.sin(np.pi * X / L) * np.cos(np.pi * Y / L))


def initial_velocity(cfg, nx, ny) -> np.ndarray:
    """The initial velocity of ``iblbm/init.py:178-191``: the analytic field at t = 0."""
    L = tg_half_width(cfg)
    XX, YY = np.meshgrid(np.arange(nx, dtype=float) - (nx - 1) / 2.0,
                         np.arange(ny, dtype=float) - (ny - 1) / 2.0)
    ux, uy = analytic_field(XX.ravel(), YY.ravel(), 0.0, cfg.tg_u0, L, tg_nu(cfg))
    U = np.zeros((nx * ny, 2))
    U[:, 0] = ux
    U[:, 1] = uy
    return U


def boundary_coefficients(cfg, nx, ny) -> np.ndarray:
    """(A, B) per boundary node, u = (A d(t), B d(t)), in the order left (j = 0..ny-1), right,
    top (i = 1..nx-2), bottom, corners (0,0), (nx-1,0), (0,ny-1), (nx-1,ny-1): the arrays of
    ``tg_analytical.py:64-136`` before their decay factor."""
    L = tg_half_width(cfg)
    X = np.arange(nx, dtype=float) - (nx - 1) / 2.0
    Y = np.arange(ny, dtype=float) - (ny - 1) / 2.0
    j = np.arange(ny)
    i = np.arange(1, nx - 1)
    parts = [_coef(np.full_like(j, X[0], dtype=float), Y[j], cfg.tg_u0, L),
             _coef(np.full_like(j, X[nx - 1], dtype=float), Y[j], cfg.tg_u0, L),
             _coef(X[i], np.full_like(i, Y[ny - 1], dtype=float), cfg.tg_u0, L),
             _coef(X[i], np.full_like(i, Y[0], dtype=float), cfg.tg_u0, L),
             _coef(X[np.array([0, nx - 1, 0, nx - 1])], Y[np.array([0, 0, ny - 1, ny - 1])],
                   cfg.tg_u0, L)]
    return np.concatenate([np.stack([a, b], 1) for a, b in parts])


def decay_table(cfg, n_steps: int, dt: float = 1.0) -> np.ndarray:
    """d(t) for t = ttt dt, ttt = 0..n_steps, each from the reference's scalar expression."""
    L, nu = tg_half_width(cfg), tg_nu(cfg)
    return np.array([np.exp(-2.0 * nu * (np.pi / L) ** 2 * (k * dt)) for k in range(n_steps + 1)],
                    dtype=np.float64)


def desired_velocity(cfg, Lx, Ly, dx, dy, nx, ny, t) -> np.ndarray:
    """``taylor_green.py:21-32``: the analytic velocity at the markers at time t."""
    Xc = Lx / dx - (nx - 1) / 2.0
    Yc = Ly / dy - (ny - 1) / 2.0
    ux, uy = analytic_field(Xc, Yc, t, cfg.tg_u0, tg_half_width(cfg), tg_nu(cfg))
    ud = np.empty((len(Xc), 2))
    ud[:, 0] = ux
    ud[:, 1] = uy
    return ud


def l2_error(U_num, U_ana) -> float:
    """``compute_l2_error`` (``iblbm/diagnostics/core.py:241-263``)."""
    diff = U_num - U_ana
    numer = np.sqrt(np.sum(diff ** 2))
    denom = np.sqrt(np.sum(U_ana ** 2))
    if denom == 0.0:
        return float("inf")
    return float(numer / denom)


def analytic_velocity(cfg, nx, ny, t) -> np.ndarray:
    """The analytic velocity (N, 2) over the whole grid at time t."""
    L = tg_half_width(cfg)
    XX, YY = np.meshgrid(np.arange(nx, dtype=float) - (nx - 1) / 2.0,
                         np.arange(ny, dtype=float) - (ny - 1) / 2.0)
    ux, uy = analytic_field(XX.ravel(), YY.ravel(), t, cfg.tg_u0, L, tg_nu(cfg))
    return np.stack([ux, uy], 1)
