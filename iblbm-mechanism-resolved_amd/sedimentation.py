"""Settling-particle runs on the GPU hot path (single disc and multi-particle DFC).

The lattice, the marker window (MDF/DF/DFC) and the collision run on the device; the rigid-
body update (Velocity Verlet with the ``none``, ``explicit_history`` or ``feng_b2`` inertia
model, ``iblbm/physics/inertia``) stays on the host in FP64 with the reference's expressions, as in
``iblbm/runtime/scenarios/sedimentation_single.py:91-293`` and ``sedimentation_multi.py:49-124``.
Per step the host uploads the moved markers and reads back the window fib (DF/MDF) or the
marker forces (DFC).  The force and torque sums are formed in NumPy's own summation order
(pairwise inside 8192-element blocks, blocks in sequence), so the trajectories carry the
reference's bits.
"""
from __future__ import annotations

import ctypes
import math
import time

import numpy as np

from . import _abi, gpu_kernels
from ._abi import check
from .init import circle_markers, initialize
from .solver import PRECISIONS, Context, make_case, pre_collide

_NP_BLOCK = 8192  # NumPy's reduction buffer: pairwise inside blocks, blocks in sequence


def np_sum_sparse(vals: np.ndarray, start: int, n_total: int) -> float:
    """``float(np.sum(v))`` of a length-``n_total`` vector that is ``vals`` on
    [start, start+len(vals)) and +0.0 elsewhere, without materialising it: only the NumPy
    reduction blocks that overlap ``vals`` are summed (with NumPy itself), in order."""
    end = start + len(vals)
    res = 0.0
    b0 = (start // _NP_BLOCK) * _NP_BLOCK
    for b in range(b0, end, _NP_BLOCK):
        blk = np.zeros(min(_NP_BLOCK, n_total - b))
        lo, hi = max(start, b), min(end, b + len(blk))
        blk[lo - b:hi - b] = vals[lo - start:hi - start]
        res += float(np.sum(blk))
    return res


class WindowSums:
    """``float(np.sum(v))`` of full-grid (N,) vectors that are zero outside one rectangular
    window, for the window's values only.  NumPy reduces a contiguous float64 vector in
    buffers of _NP_BLOCK elements aligned to element 0 (pairwise inside a buffer, buffers
    added in order); the window's values are scattered into zeroed copies of the buffers
    they fall in, every buffer of every vector is reduced by one NumPy call along its last
    axis (the same pairwise reduction per buffer), and the buffer sums are added in order.
    The scatter indices are kept per window geometry."""

    def __init__(self, nx: int, n_total: int):
        self.nx, self.N = int(nx), int(n_total)
        self._key = None

    def _layout(self, box):
        if box == self._key:
            return
        x0, y0, bw, bh = box
        ii = np.arange(bw)
        flat = ((y0 + np.arange(bh))[:, None] * self.nx + (x0 + ii)[None, :]).ravel()
        b0 = flat[0] // _NP_BLOCK
        b1 = flat[-1] // _NP_BLOCK
        self.nb = int(b1 - b0 + 1)
        self.rel = flat - b0 * _NP_BLOCK  # offsets into the (nb * _NP_BLOCK) buffers
        # a buffer cut short by the end of the vector is reduced on its own, at its length
        self.last_len = min(_NP_BLOCK, self.N - int(b1) * _NP_BLOCK)
        # zeroed once per geometry: every call writes the same window positions only
        self.buf = np.zeros((3, self.nb * _NP_BLOCK))
        self._key = box

    def sums(self, box, *vals):
        """One NumPy-order sum per window-shaped value array in vals (at most three)."""
        self._layout(box)
        k = len(vals)
        for q, v in enumerate(vals):
            self.buf[q, self.rel] = v
        buf = self.buf[:k].reshape(k, self.nb, _NP_BLOCK)
        if self.last_len == _NP_BLOCK:
            part = np.sum(buf, axis=2)
        else:
            part = np.empty((k, self.nb))
            part[:, :-1] = np.sum(buf[:, :-1, :], axis=2)
            for q in range(k):
                part[q, -1] = np.sum(buf[q, -1, :self.last_len])
        out = []
        for q in range(k):
            res = 0.0
            for x in part[q]:
                res += float(x)
            out.append(res)
        return out


def compute_net_gravity(rho_ratio, r_lattice, g_lattice, gravity_direction="down"):
    """``iblbm/physics/gravity.py:16-34``."""
    mag = (rho_ratio - 1.0) * (np.pi * r_lattice**2) * g_lattice
    if gravity_direction == "right":
        return np.array([mag, 0.0])
    return np.array([0.0, -mag])


def _history_dv(model, vel, vel_prev, mass, r_lattice):
    if model == "explicit_history":  # inertia/explicit_history.py:41-54
        m_fluid = np.pi * r_lattice**2
        return (m_fluid / mass) * (vel - vel_prev)
    return None


def verlet_half(model, pos, vel, vel_prev, force, mass, rho_ratio, r_lat, g, dt, dx, gdir):
    """``velocity_verlet_half`` of the ``none`` / ``explicit_history`` inertia models
    (``feng_b2`` integrates as ``none``, ``inertia/feng_b2.py:43-48``)."""
    F_total = force + compute_net_gravity(rho_ratio, r_lat, g, gdir)
    hdv = _history_dv(model, vel, vel_prev, mass, r_lat)
    if hdv is None:
        vel_half = vel + 0.5 * dt * F_total / mass
    else:
        vel_half = vel + 0.5 * dt * F_total / mass + 0.5 * hdv
    return pos + dt * vel_half * dx, vel_half


def verlet_full(model, vel_half, vel, vel_prev, force, mass, rho_ratio, r_lat, g, dt, gdir):
    F_total = force + compute_net_gravity(rho_ratio, r_lat, g, gdir)
    hdv = _history_dv(model, vel, vel_prev, mass, r_lat)
    if hdv is None:
        return vel_half + 0.5 * dt * F_total / mass
    return vel_half + 0.5 * dt * F_total / mass + 0.5 * hdv


def euler_explicit(model, pos, vel, vel_prev, force, mass, rho_ratio, r_lat, g, dt, dx, scheme,
                   gdir):
    """``euler_explicit`` of the ``none`` / ``explicit_history`` inertia models
    (``inertia/none.py:48-59``, ``inertia/explicit_history.py:95-114``; ``feng_b2`` integrates
    as ``none``) with the position update ``euler_explicit_position_update``
    (``rigid_body.py:85-98``): x += dt u dx with u the new velocity or the mean of the old and
    the new (``trapezoidal``)."""
    F_total = force + compute_net_gravity(rho_ratio, r_lat, g, gdir)
    hdv = _history_dv(model, vel, vel_prev, mass, r_lat)
    if hdv is None:
        vel_new = vel + dt * F_total / mass
    else:
        vel_new = vel + dt * F_total / mass + hdv
    if scheme == "new_velocity":
        pos_new = pos + dt * vel_new * dx
    elif scheme == "trapezoidal":
        pos_new = pos + 0.5 * dt * (vel + vel_new) * dx
    else:
        raise ValueError(f"sedimentation_euler_update_scheme={scheme!r}: 'new_velocity' or "
                         "'trapezoidal'")
    return pos_new, vel_new


def euler_rotation(omega, torque, I, dt, dl_int, scheme):
    """``euler_explicit_rotation`` (``rigid_body.py:101-121``): omega += (dL_int + dt T) / I,
    the angle advanced by dt times the new rate or the mean of the old and the new."""
    omega_new = omega + (dl_int + dt * torque) / I
    if scheme == "new_velocity":
        return omega_new, dt * omega_new
    if scheme == "trapezoidal":
        return omega_new, 0.5 * dt * (omega + omega_new)
    raise ValueError(f"sedimentation_euler_update_scheme={scheme!r}: 'new_velocity' or "
                     "'trapezoidal'")


def l_int_box(nx, ny, dx, cx, cy, r_domain):
    """The node box [i_lo, i_hi) x [j_lo, j_hi) around a disc that ``compute_l_int_z``
    reads (``diagnostics/internal_flow.py:237-249``: the rounded centre +- int(r) + 2,
    clipped to the grid)."""
    ci, cj = cx / dx, cy / dx
    r_idx = int(r_domain / dx) + 2
    ci_int, cj_int = int(round(ci)), int(round(cj))
    return (max(0, ci_int - r_idx), min(nx, ci_int + r_idx + 1),
            max(0, cj_int - r_idx), min(ny, cj_int + r_idx + 1))


def l_int_z(sub_ro, sub_U, i_lo, j_lo, dx, cx, cy, r_domain):
    """Internal angular momentum of the fluid inside a disc, sum over the nodes within r of
    the centre of (r_x rho u_y - r_y rho u_x) (``compute_l_int_z``,
    ``diagnostics/internal_flow.py:220-258``), from rho and U on the box ``l_int_box`` gives
    ((h, w) and (h, w, 2), box-major) in the same NumPy operations and order."""
    ci, cj = cx / dx, cy / dx
    r_lat = r_domain / dx
    h, w = sub_ro.shape
    x_lat = np.arange(i_lo, i_lo + w, dtype=float) - ci
    y_lat = np.arange(j_lo, j_lo + h, dtype=float) - cj
    RX, RY = np.meshgrid(x_lat, y_lat)
    mask = RX**2 + RY**2 <= r_lat**2
    sub_Eux, sub_Euy = sub_U[:, :, 0], sub_U[:, :, 1]
    return float(np.sum(RX[mask] * sub_ro[mask] * sub_Euy[mask] -
                        RY[mask] * sub_ro[mask] * sub_Eux[mask]))


def euler_step_particles(particles, forces, cfg, st, g, state_box):
    """The explicit Euler step of every particle from its (force, torque), then its moved
    markers and desired velocities: the single-particle ``post_ibm``
    (``sedimentation_single.py:187-262``) and ``_multi_particle_euler_step``
    (``sedimentation_multi.py:127-202``).  state_box(i_lo, j_lo, w, h) returns rho and U of the
    step just run on that box."""
    scheme = cfg.sedimentation_euler_update_scheme
    for p, (F, T) in zip(particles, forces):
        pos_new, vel_new = euler_explicit(cfg.settling_inertia_model, p.pos, p.vel, p.vel_prev,
                                          F, p.mass, p.rho_ratio, p.lattice_r, g, st.dt, st.dx,
                                          scheme, cfg.gravity_direction)
        omega_new = 0.0
        if cfg.enable_rotation:
            cx, cy = float(p.pos[0]), float(p.pos[1])  # the position before this update
            i_lo, i_hi, j_lo, j_hi = l_int_box(st.nx, st.ny, st.dx, cx, cy, p.r_domain)
            ro, U = state_box(i_lo, j_lo, i_hi - i_lo, j_hi - j_lo)
            l_now = l_int_z(ro, U, i_lo, j_lo, st.dx, cx, cy, p.r_domain)
            dl_int = l_now - p.l_int_prev
            p.l_int_prev = l_now
            omega_new, angle_inc = euler_rotation(p.omega, T, p.I, st.dt, dl_int, scheme)
            p.angle += angle_inc
            p.torque = T
        vel_new = post_correction(cfg.settling_inertia_model, vel_new, p.vel, p.vel_prev,
                                  p.rho_ratio)
        if cfg.enable_rotation:
            p.omega = omega_new
        p.vel_prev = p.vel.copy()
        p.vel = vel_new
        p.pos = pos_new
        p.force = F
        p.Lx, p.Ly = update_markers(p.Lx_c, p.Ly_c, p.cx_init, p.cy_init, pos_new,
                                    angle=p.angle if cfg.enable_rotation else 0.0)
        if cfg.enable_rotation:
            p.desired = desired_velocity(vel_new, p.Lb, omega=omega_new, Lx=p.Lx, Ly=p.Ly,
                                         cx=float(pos_new[0]), cy=float(pos_new[1]), dx=st.dx)
        else:
            p.desired = desired_velocity(vel_new, p.Lb)


def post_correction(model, vel_new, vel, vel_prev, rho_ratio):
    """The inertia model's correction after the full step: ``feng_b2`` adds
    (rho_f / rho_s) (U^n - U^{n-1}) (``inertia/feng_b2.py:33-48``); the others return the
    velocity unchanged (``inertia/none.py:61-62``)."""
    if model == "feng_b2":
        return vel_new + (1.0 / rho_ratio) * (vel - vel_prev)
    return vel_new


def update_markers(Lx_c, Ly_c, cx_init, cy_init, pos_new, angle=0.0):
    """``iblbm/physics/markers.py:23-38``."""
    cos_a = np.cos(angle)
    sin_a = np.sin(angle)
    rx = Lx_c - cx_init
    ry = Ly_c - cy_init
    return cos_a * rx - sin_a * ry + pos_new[0], sin_a * rx + cos_a * ry + pos_new[1]


def desired_velocity(vel, Lb, omega=0.0, Lx=None, Ly=None, cx=0.0, cy=0.0, dx=1.0):
    """``iblbm/physics/markers.py:41-57``."""
    d = np.empty((Lb, 2))
    d[:, 0] = float(vel[0])
    d[:, 1] = float(vel[1])
    if omega != 0.0 and Lx is not None:
        d[:, 0] += -omega * (Ly - cy) / dx
        d[:, 1] += omega * (Lx - cx) / dx
    return d


def record_state(pos, vel, t, d_lattice, g, rho_ratio, y0, dx):
    """``iblbm/diagnostics/internal_flow.py:156-218``."""
    u_g = np.sqrt(abs(rho_ratio - 1.0) * g * d_lattice)
    D_domain = d_lattice * dx
    if u_g < 1e-15:
        return {"step": int(t), "t_star": 0.0, "y_star": 0.0, "vy_star": 0.0, "vx_star": 0.0,
                "x": float(pos[0]), "y": float(pos[1]), "vx": float(vel[0]),
                "vy": float(vel[1])}
    return {"step": int(t), "t_star": float(t * u_g / d_lattice),
            "y_star": float((y0 - pos[1]) / D_domain), "vy_star": float(-vel[1] / u_g),
            "vx_star": float(vel[0] / u_g), "x": float(pos[0]), "y": float(pos[1]),
            "vx": float(vel[0]), "vy": float(vel[1])}


class Particle:
    def __init__(self, cx, cy, rho_ratio, lattice_r, r_domain, dx, spacing, retraction, Lx=None,
                 Ly=None, Lb=None, Larea=None):
        if Lx is None:
            Lx, Ly, Lb, Larea = circle_markers(cx, cy, r_domain, lattice_r, dx, spacing, retraction)
        self.pos = np.array([cx, cy], dtype=float)
        self.vel = np.array([0.0, 0.0], dtype=float)
        self.vel_prev = np.array([0.0, 0.0], dtype=float)
        self.force = np.array([0.0, 0.0], dtype=float)
        self.rho_ratio = rho_ratio
        self.lattice_r = lattice_r
        self.r_domain = r_domain
        self.mass = rho_ratio * np.pi * lattice_r**2  # physics/sedimentation.py:126-127
        self.I = 0.5 * self.mass * lattice_r**2
        self.omega = 0.0
        self.torque = 0.0
        self.angle = 0.0
        self.Lx_c, self.Ly_c = Lx.copy(), Ly.copy()
        self.Lx, self.Ly = Lx, Ly
        self.Lb, self.Larea = Lb, Larea
        self.desired = np.zeros((Lb, 2))
        self.cx_init, self.cy_init = cx, cy
        self.vel_half = None
        self.omega_half = 0.0
        self.l_int_prev = 0.0  # explicit Euler: the fluid's angular momentum inside the disc


def make_particles(cfg, st):
    """The particles of a settling run: ``particles_config`` (``sedimentation_multi.py``) or
    the one disc at ``cylinder_center`` with the initial marker ring."""
    if cfg.particles_config is not None:
        parts = [Particle(pc["center"][0], pc["center"][1], pc["rho_ratio"], st.lattice_r, st.r,
                          st.dx, cfg.marker_spacing_factor, cfg.retraction_dx)
                 for pc in cfg.particles_config]
        if len(parts) > 8:
            raise ValueError("the GPU window holds at most 8 particles")
        larea = parts[0].Larea
        if cfg.ibm_method == "DFC":  # strategies/dfc.py:56-58
            for p in parts:
                if abs(float(p.Larea) - larea) > 1e-12:
                    raise ValueError("multi-particle DFC needs one Larea for every particle")
        return parts
    cx, cy = cfg.cylinder_center
    return [Particle(cx, cy, cfg.rho_ratio, st.lattice_r, st.r, st.dx, cfg.marker_spacing_factor,
                     cfg.retraction_dx, Lx=st.Lx, Ly=st.Ly, Lb=st.Lb, Larea=st.Larea)]


def half_step_particles(particles, cfg, st, g, multi):
    """The Verlet half step of every particle, its moved markers and desired velocities:
    ``_multi_particle_half_step`` (``sedimentation_multi.py:49-94``) and the single-particle
    ``pre_step`` (``sedimentation_single.py:91-160``)."""
    for p in particles:
        pos_new, vel_half = verlet_half(cfg.settling_inertia_model, p.pos, p.vel, p.vel_prev,
                                        p.force, p.mass, p.rho_ratio, p.lattice_r, g,
                                        st.dt, st.dx, cfg.gravity_direction)
        p.pos = pos_new
        if cfg.enable_rotation:  # rigid_body.py:65-74
            p.omega_half = p.omega + 0.5 * st.dt * p.torque / p.I
            p.angle += st.dt * p.omega_half
        else:
            p.omega_half = 0.0
        p.Lx, p.Ly = update_markers(p.Lx_c, p.Ly_c, p.cx_init, p.cy_init, pos_new,
                                    angle=p.angle if cfg.enable_rotation else 0.0)
        if cfg.rotation_coupling == "semi_implicit" and cfg.enable_rotation:
            omega = p.omega_half if not multi else p.omega
            p.desired = desired_velocity(vel_half, p.Lb, omega=omega, Lx=p.Lx, Ly=p.Ly,
                                         cx=float(p.pos[0]), cy=float(p.pos[1]), dx=st.dx)
        else:
            p.desired = desired_velocity(vel_half, p.Lb)
        p.vel_half = vel_half


def full_step_particles(particles, forces, cfg, st, g):
    """The Verlet full step of every particle from its (force, torque):
    ``_multi_particle_full_step`` (``sedimentation_multi.py:97-124``) and the single-particle
    ``post_ibm`` (``sedimentation_single.py:200-293``)."""
    for p, (F, T) in zip(particles, forces):
        vel_new = verlet_full(cfg.settling_inertia_model, p.vel_half, p.vel, p.vel_prev, F,
                              p.mass, p.rho_ratio, p.lattice_r, g, st.dt,
                              cfg.gravity_direction)
        if cfg.enable_rotation:  # rigid_body.py:77-82
            p.omega = p.omega_half + 0.5 * st.dt * T / p.I
            p.torque = T
        vel_new = post_correction(cfg.settling_inertia_model, vel_new, p.vel, p.vel_prev,
                                  p.rho_ratio)
        p.vel_prev = p.vel.copy()
        p.vel = vel_new
        p.force = F


def marker_forces_to_particles(particles, q, cfg, st):
    """(force, torque) per particle from the marker forces of its ring (rings concatenated in
    particle order): DFC sums q (``strategies/dfc.py:129-142``), several-particle DF sums
    -Lf * Larea (``forcing.py:263-275``)."""
    out = []
    c = 0
    for p in particles:
        fq = q[c:c + p.Lb]
        c += p.Lb
        if cfg.ibm_method == "DFC":
            F = np.array([float(fq[:, 0].sum()), float(fq[:, 1].sum())])
            T = 0.0
            if cfg.enable_rotation:
                rx = (p.Lx - float(p.pos[0])) / st.dx
                ry = (p.Ly - float(p.pos[1])) / st.dy
                T = float(np.sum(rx * fq[:, 1] - ry * fq[:, 0]))
        else:
            Lfx, Lfy = fq[:, 0], fq[:, 1]
            la = float(p.Larea)
            F = np.array([-float(np.sum(Lfx * la)), -float(np.sum(Lfy * la))])
            T = 0.0
            if cfg.enable_rotation:
                rx = (p.Lx - float(p.pos[0])) / st.dx
                ry = (p.Ly - float(p.pos[1])) / st.dy
                T = -float(np.sum((rx * Lfy - ry * Lfx) * la))
        out.append((F, T))
    return out


class Sedimentation:
    """One settling-particle instance on the device, stepped one time step at a time."""

    def __init__(self, cfg, device=0, precision="fp64", ctx=None, f0_cache=None,
                 mdf_omega=None):
        """mdf_omega: the relaxation factor the MDF window starts from each step, in place of
        the one derived from the A-norm estimate (several particles: min(0.8 omega, 0.8),
        ``multi_particle.py:85-91``); the reference takes it from its cached estimate
        ``state._mdf_omega``, which a caller may preset."""
        if cfg.motion_type != "sedimentation":
            raise ValueError("Sedimentation needs motion_type='sedimentation'")
        if cfg.settling_inertia_model not in ("none", "explicit_history", "feng_b2"):
            raise ValueError(f"settling_inertia_model={cfg.settling_inertia_model!r} is outside "
                             "the GPU hot path (supported: 'none', 'explicit_history', "
                             "'feng_b2')")
        if cfg.time_integrator not in ("verlet", "euler_explicit"):
            raise ValueError(f"time_integrator={cfg.time_integrator!r}: 'verlet' or "
                             "'euler_explicit'")
        if cfg.rotation_coupling == "iterative" and cfg.enable_rotation:
            raise ValueError("rotation_coupling='iterative' is outside the GPU hot path")
        if cfg.prescribed_velocity is not None and cfg.particles_config is not None:
            # the reference's particle-list runtime has no prescribed-velocity branch
            raise ValueError("prescribed_velocity applies to the single-particle run")
        if cfg.gravity <= 0.0:
            raise ValueError(f"sedimentation needs gravity > 0 (got {cfg.gravity})")
        if cfg.rho_ratio <= 1.0:
            raise ValueError(f"sedimentation needs rho_ratio > 1.0 (got {cfg.rho_ratio})")
        self.cfg = cfg
        # the pre-collided rest state depends on the lattice and collision only, so members of
        # an ensemble on one grid share it (f0_cache) and keep no full-grid host fields
        st = initialize(cfg, fields=False)
        key = (st.nx, st.ny, st.tau, st.tau_minus, cfg.collision_model,
               None if st.S_diag is None else tuple(float(v) for v in st.S_diag),
               bool(cfg.incompressible_lbgk))
        f0 = None if f0_cache is None else f0_cache.get(key)
        if f0 is None:
            f0 = pre_collide(cfg, initialize(cfg))
            if f0_cache is not None:
                f0_cache[key] = f0
        self.st = st
        self.g = cfg.gravity
        self.multi = cfg.particles_config is not None
        self.particles = make_particles(cfg, st)
        self.ctx = ctx or Context(device)
        case = make_case(cfg, st, PRECISIONS[precision])
        case.fixed_body = 0
        case.lb_capacity = sum(p.Lb for p in self.particles)
        case.box_capacity = st.nx * st.ny
        self.omega = None
        if cfg.ibm_method == "MDF":  # strategies/mdf.py:27-41 (the ring at cylinder_center)
            om = gpu_kernels.estimate_optimal_omega_from_type(st.Lx, st.Ly, cfg.delta_type,
                                                              st.Larea, st.dx)
            if cfg.enable_rotation:
                om = min(float(om), 1.0)
            self.omega = om
            # several particles iterate coupled from min(0.8 omega, 0.8)
            # (multi_particle.py:32-36, :85-86); a one-element particle list runs the
            # single-particle MDF with the estimate itself (multi_particle.py:432-448)
            case.mdf_omega = min(om * 0.8, 0.8) if len(self.particles) > 1 else om
            if mdf_omega is not None:
                case.mdf_omega = float(mdf_omega)
        self.k = self.ctx.add(case)
        if self.multi and cfg.ibm_method in ("DF", "MDF"):
            self.ctx.set_particles(self.k, [p.Lb for p in self.particles],
                                   [float(p.Larea) for p in self.particles])
        self._push_markers()
        self.ctx.upload_f(self.k, f0)
        self.ttt = 0
        self.history = []
        self._mdf_read = None
        self.d_lattice = 2.0 * st.lattice_r
        self.termination_reason = "max_steps"
        self._wsums = None

    @property
    def predictor(self):
        """The learned MDF force predictor of this instance (``iblbm_amd.predictor``)."""
        from .predictor import Predictor
        return Predictor(self.ctx, self.k)

    def snapshot(self):
        """Device populations and host particle state, for ``restore``."""
        import copy
        return self.ctx.download_f(self.k), copy.deepcopy(self.particles), self.ttt

    def restore(self, snap):
        import copy
        f, parts, ttt = snap
        self.particles = copy.deepcopy(parts)
        self.ttt = ttt
        self._push_markers()
        self.ctx.upload_f(self.k, f)

    def _push_markers(self):
        """The moved rings, desired velocities and torque centre (particle 0) in one call
        with one synchronisation (``iblbm_set_markers_batch`` for this one instance)."""
        Lx = np.concatenate([p.Lx for p in self.particles])
        Ly = np.concatenate([p.Ly for p in self.particles])
        ud = np.concatenate([p.desired for p in self.particles])
        p0 = self.particles[0]
        self.ctx.set_markers_batch([self.k], [len(Lx)], Lx, Ly, ud, [float(p0.Larea)],
                                   center=[float(p0.pos[0]), float(p0.pos[1])])

    # -- the reference's pre_step / post_ibm hooks ------------------------------------------
    def _half_step(self):
        half_step_particles(self.particles, self.cfg, self.st, self.g, self.multi)

    def _forces(self):
        """Hydrodynamic force and torque per particle (strategies/*.py extract_force/torque)."""
        st, cfg = self.st, self.cfg
        out = []
        if cfg.ibm_method == "DFC":
            Lbt = sum(p.Lb for p in self.particles)
            return marker_forces_to_particles(self.particles,
                                              self.ctx.download_dfc_force(self.k, Lbt), cfg, st)
        if self.multi and cfg.ibm_method == "DF":  # forcing.py:263-275, per particle
            Lbt = sum(p.Lb for p in self.particles)
            return marker_forces_to_particles(self.particles,
                                              self.ctx.download_marker_force(self.k, Lbt), cfg, st)
        if self.multi and len(self.particles) > 1:
            # coupled MDF: each particle's own field (multi_particle.py:302-316); one particle
            # runs the single-particle MDF, whose force and torque are the window fib's sums
            # (multi_particle.py:432-448; no published run holds this case)
            for i, p in enumerate(self.particles):
                out.append(self._window_force(*self.ctx.download_particle_fib(self.k, i), p))
            return out
        out.append(self._window_force(*self.ctx.download_window(self.k), self.particles[0]))
        return out

    def _window_force(self, fx, fy, box, p):
        """-(grid sums) of a window fib and its torque about particle p, in NumPy's order
        over the full (N,) / (ny, nx) arrays (strategies/mdf.py:86-101).  Outside the
        window both force components are zero, and so is the torque term there (a signed
        zero, which changes no sum that is not itself zero)."""
        cfg, st = self.cfg, self.st
        x0, y0, bw, bh = box
        if self._wsums is None:
            self._wsums = WindowSums(st.nx, st.nodenums)
        fx = np.asarray(fx, dtype=np.float64)
        fy = np.asarray(fy, dtype=np.float64)
        if not cfg.enable_rotation:
            sx, sy = self._wsums.sums(box, fx, fy)
            return -np.array([sx, sy]), 0.0
        # strategies/mdf.py:94-101: (i - cx) fy - (j - cy) fx over the grid
        cx_lat = float(p.pos[0]) / st.dx
        cy_lat = float(p.pos[1]) / st.dy
        II = np.tile(np.arange(x0, x0 + bw, dtype=float), bh)
        JJ = np.repeat(np.arange(y0, y0 + bh, dtype=float), bw)
        term = (II - cx_lat) * fy - (JJ - cy_lat) * fx
        sx, sy, tq = self._wsums.sums(box, fx, fy, term)
        return -np.array([sx, sy]), -tq

    def _full_step(self, forces):
        full_step_particles(self.particles, forces, self.cfg, self.st, self.g)

    def _prescribed_step(self, ci):
        """A disc moved at a prescribed velocity (``sedimentation_single.py:92-98``,
        ``:176-181``): it moves by dt u dx before the IBM, its ring follows without rotation,
        the desired velocity is u, and after the step its velocity is u and its force the
        hydrodynamic force (no torque, no inertia model)."""
        st, p = self.st, self.particles[0]
        pv = np.array(self.cfg.prescribed_velocity, dtype=float)
        p.pos = p.pos + st.dt * pv * st.dx
        p.Lx, p.Ly = update_markers(p.Lx_c, p.Ly_c, p.cx_init, p.cy_init, p.pos)
        p.desired = desired_velocity(pv, p.Lb)
        self._push_markers()
        self.ctx.step(1, diag_every=1 if self.ttt % ci == 0 else 0)
        F, _ = self._forces()[0]
        p.vel = np.array(self.cfg.prescribed_velocity, dtype=float)
        p.force = F

    def step(self):
        self.ttt += 1
        if self.cfg.prescribed_velocity is not None:
            self._prescribed_step(int(self.cfg.check_interval))
            self._mdf_read = None
            return
        euler = self.cfg.time_integrator == "euler_explicit"
        if not euler:  # Verlet: half step before the IBM (explicit Euler moves after it)
            self._half_step()
            self._push_markers()
        ci = int(self.cfg.check_interval)
        self.ctx.step(1, diag_every=1 if self.ttt % ci == 0 else 0)
        if euler:
            euler_step_particles(self.particles, self._forces(), self.cfg, self.st, self.g,
                                 lambda x0, y0, w, h: self.ctx.download_state_box(
                                     self.k, x0, y0, w, h))
            self._push_markers()  # the markers of the next step
        else:
            self._full_step(self._forces())
        self._mdf_read = None  # the last step's MDF statistics are read when asked for

    @property
    def mdf_stats(self):
        """MDF statistics of the last step (``mdf.py:224-229``), or None for DF / DFC."""
        if self.cfg.ibm_method != "MDF" or self.ttt == 0:
            return None
        if self._mdf_read is None:
            self._mdf_read = self.ctx.mdf_stats(self.k)
        return self._mdf_read

    def record(self):
        cfg, st = self.cfg, self.st
        recs = [record_state(p.pos, p.vel, self.ttt, self.d_lattice, self.g, p.rho_ratio,
                             p.cy_init, st.dx) for p in self.particles]
        return recs

    def should_stop(self):
        """``sedimentation_single.py:339-351`` / ``sedimentation_multi.py:298-313``."""
        vy_star = None
        if not self.multi:
            vy_star = self.record()[0]["vy_star"]
        return stop_reason(self.cfg, self.particles, self.st.dx, self.multi, vy_star)


def _stop_rule(pos, r_domain, cfg):
    """``sedimentation_common.py:14-34``: "contact" / "offset" of the leading particle."""
    if cfg.sedimentation_stop_at_contact:
        if cfg.gravity_direction == "right":
            if float(pos[0]) >= cfg.xmax - r_domain:
                return "contact"
        elif float(pos[1]) <= r_domain:
            return "contact"
    if cfg.sedimentation_stop_offset_d > 0.0:
        d_dom = cfg.cylinder_D_ratio
        if cfg.gravity_direction == "right":
            if float(pos[0]) >= cfg.xmax - cfg.sedimentation_stop_offset_d * d_dom:
                return "offset"
        elif float(pos[1]) <= cfg.sedimentation_stop_offset_d * d_dom:
            return "offset"
    return None


def _near_wall(pos, r_domain, cfg, dx, safety=2.0):
    """``sedimentation_common.py:37-66``: True where the reference returns a warning (the
    particle within ``safety*dx`` of a wall other than the stopping wall)."""
    margin = safety * dx
    x, y = float(pos[0]), float(pos[1])
    terminal = cfg.sedimentation_stop_at_contact or cfg.sedimentation_stop_offset_d > 0.0
    if cfg.gravity_direction == "right":
        return (y - r_domain < margin or y + r_domain > cfg.ymax - margin or
                x - r_domain < margin or (not terminal and x + r_domain > cfg.xmax - margin))
    return ((not terminal and y - r_domain < margin) or y + r_domain > cfg.ymax - margin or
            x - r_domain < margin or x + r_domain > cfg.xmax - margin)


def stop_reason(cfg, particles, dx, multi, vy_star=None):
    """The settling stop rules, in the reference's order: "nan", then "contact" / "offset" of
    particle 0, then "domain_bounds" (any particle near a wall), else None.

    One particle (``sedimentation_single.py:339-351``): "nan" when the record's ``vy_star``
    (``-v_y / u_g``, 0 when ``u_g < 1e-15``) is not finite, and only particle 0 is checked
    against the walls.  Several (``sedimentation_multi.py:298-313``): "nan" when particle 0's
    velocity along gravity is not finite, and every particle is checked against the walls.
    """
    p0 = particles[0]
    if multi:
        v = float(p0.vel[0]) if cfg.gravity_direction == "right" else float(p0.vel[1])
    else:
        v = float(vy_star)
    if not math.isfinite(v):
        return "nan"
    reason = _stop_rule(p0.pos, p0.r_domain, cfg)
    if reason is not None:
        return reason
    for p in (particles if multi else particles[:1]):
        if _near_wall(p.pos, p.r_domain, cfg, dx):
            return "domain_bounds"
    return None


def run(cfg, verbose: bool = True, callback=None, save_dir=None, device: int = 0,
        precision: str = "fp64") -> dict:
    """``iblbm.run`` for sedimentation: result keys of the reference plus the history."""
    sim = Sedimentation(cfg, device=device, precision=precision)
    mon = None
    if save_dir:  # status.json + sedimentation_history.json in the reference's format
        from .monitor import SedimentationMonitor
        mon = SedimentationMonitor(save_dir, cfg)
    ci = int(cfg.check_interval)
    t0 = time.time()
    while sim.ttt < cfg.max_steps:
        sim.step()
        if sim.ttt % ci == 0:
            recs = sim.record()
            sim.history.append(recs if sim.multi else recs[0])
            if mon is not None:
                mon(sim.ttt, sim.particles, sim.mdf_stats, elapsed=time.time() - t0)
            if verbose:
                r0 = recs[0]
                print(f"step {sim.ttt:>7d} | y*={r0['y_star']:.4f} vy*={r0['vy_star']:.6f}",
                      flush=True)
            if callback is not None:
                callback(step=sim.ttt, Cd=0.0, Cl=0.0, error=1.0, state=sim, elapsed=0.0,
                         converged=False)
            reason = sim.should_stop()
            if reason is not None:
                sim.termination_reason = reason
                break
    sim.ctx.sync()
    final_step = stop_final_step(sim.ttt, sim.termination_reason)
    if mon is not None:
        mon.finalize(final_step, sim.termination_reason, tau=sim.st.tau,
                     rho_ratios=[float(p.rho_ratio) for p in sim.particles])
    return {"Cd_history": np.array([]), "Cl_history": np.array([]), "converged": False,
            "final_step": final_step, "termination_reason": sim.termination_reason,
            "sedimentation_history": sim.history, "sim": sim}


def stop_final_step(ttt: int, reason) -> int:
    """The ``final_step`` the reference reports: a stop rule breaks out of its loop before
    the step counter advances, and ``final_step`` is the counter minus one
    (``iblbm/runtime/loop.py:106-109``, ``:142-148``), so a run stopped at check step ``ttt``
    reports ``ttt - 1``; a run that reaches ``max_steps`` reports ``max_steps``."""
    return int(ttt) - 1 if reason not in (None, "max_steps") else int(ttt)


# Channel presets of the single-particle settling benchmark (``scenarios/sedimentation.py:20-
# 33``): height in channel widths and the initial centre, in units of the width.
DOMAIN_PRESETS = {"glowinski2001": {"ymax": 3.0, "cylinder_center": (0.5, 2.0)},
                  "tall60D": {"ymax": 7.5, "cylinder_center": (0.5, 6.5)},
                  "tall80D": {"ymax": 10.0, "cylinder_center": (0.5, 9.0)}}


def make_sedimentation_config(rho_ratio=1.01, ibm_method="DF", delta_type="peskin4pt",
                              collision_model="BGK", NN=None, Ma_max=0.08, tau_min=0.52,
                              case_name="glowinski2001", center_x_offset_dx=0.0,
                              diagnostics_interval=0, retraction_dx=0.0, nu_phys_override=None,
                              enable_rotation=False, rotation_coupling="indirect",
                              mdf_iterations=20, time_integrator="verlet",
                              incompressible_lbgk=False, settling_inertia_model="none",
                              sedimentation_stop_at_contact=False,
                              sedimentation_stop_offset_d=0.0,
                              sedimentation_reference_basis="particle_basis"):
    """The single-disc settling configuration of ``make_sedimentation_config``
    (``scenarios/sedimentation.py:36-249``): a disc of diameter d = W / 8 in a channel of
    width W = 2 cm, g = 980 cm/s^2, nu = 0.1 cm^2/s at a density ratio of 1.25 and 0.01
    otherwise (or ``nu_phys_override``).  The lattice follows from a Mach bound:

    * Archimedes number Ar = |rho_ratio - 1| g d^3 / nu^2;
    * the smallest lattice diameter D_min = 4 sqrt(3 Ar) nu_min / Ma_max with
      nu_min = (tau_min - 0.5) / 3, hence NN_min = ceil(D_min / (d / W)) + 1; without NN the
      first of 81, 161, 321, 641, 1281 at least NN_min;
    * nu_lat = max(0.8 Ma_max D / sqrt(3 Ar), nu_min), g_lat = Ar nu_lat^2 / (|rho - 1| D^3),
      the terminal-speed estimate u_t = sqrt(|rho - 1| g_lat D);
    * max_steps = int(2 (ymax / dx) / u_t): twice the channel height at u_t.

    Every expression keeps the reference's operand order, so the floats (g_lat above all) are
    the ones its result files record."""
    from .config import SimConfig
    W_phys, d_phys, g_phys = 2.0, 0.25, 980.0
    nu_phys = 0.1 if abs(rho_ratio - 1.25) < 1e-6 else 0.01
    if nu_phys_override is not None:
        nu_phys = float(nu_phys_override)
    if case_name not in DOMAIN_PRESETS:
        raise ValueError(f"case_name={case_name!r} (expected one of {sorted(DOMAIN_PRESETS)})")
    preset = DOMAIN_PRESETS[case_name]
    xmax, ymax = 1.0, preset["ymax"]
    D_ratio = d_phys / W_phys
    delta_rho = abs(rho_ratio - 1.0)
    Ar = delta_rho * g_phys * d_phys**3 / nu_phys**2
    nu_lat_min = (tau_min - 0.5) / 3.0
    if delta_rho > 1e-15 and Ar > 0:
        D_lat_min = 4.0 * np.sqrt(3.0 * Ar) * nu_lat_min / Ma_max
        NN_min = int(np.ceil(D_lat_min / D_ratio)) + 1
    else:
        NN_min = 81
    if NN is None:
        NN = next((c for c in (81, 161, 321, 641, 1281) if c >= NN_min), NN_min)
    dx = 1.0 / (NN - 1)
    lattice_D = D_ratio * (NN - 1)
    if delta_rho > 1e-15 and Ar > 0:
        nu_lat = max(Ma_max * lattice_D / np.sqrt(3.0 * Ar) * 0.8, nu_lat_min)
    else:
        nu_lat = 0.05
    g_lattice = Ar * nu_lat**2 / (delta_rho * lattice_D**3) if delta_rho > 1e-15 else 1e-6
    u_t_est = np.sqrt(delta_rho * g_lattice * lattice_D) if delta_rho > 0 else 0.0
    lattice_u = max(u_t_est, 0.01)
    Re_input = lattice_u * lattice_D / nu_lat
    max_steps = int(2.0 * (ymax / dx) / u_t_est) if u_t_est > 1e-10 else 500_000
    cx_base, cy = preset["cylinder_center"]
    return SimConfig(
        Re=Re_input, NN=NN, xmax=xmax, ymax=ymax, cylinder_center=(cx_base + center_x_offset_dx
                                                                   * dx, cy),
        cylinder_D_ratio=D_ratio, lattice_u=lattice_u, inflow_u=0.0,
        bc_type="settling_channel", delta_type=delta_type, ibm_method=ibm_method,
        collision_model=collision_model, mdf_iterations=mdf_iterations, max_steps=max_steps,
        use_convergence=False, check_interval=100, motion_type="sedimentation",
        rho_ratio=rho_ratio, gravity=g_lattice, enable_rotation=enable_rotation,
        retraction_dx=retraction_dx, rotation_coupling=rotation_coupling,
        marker_spacing_factor=1.0, nx_formula="standard",
        diagnostics_interval=diagnostics_interval, time_integrator=time_integrator,
        incompressible_lbgk=incompressible_lbgk, settling_inertia_model=settling_inertia_model,
        sedimentation_stop_at_contact=sedimentation_stop_at_contact,
        sedimentation_stop_offset_d=sedimentation_stop_offset_d,
        sedimentation_reference_basis=sedimentation_reference_basis)


def make_two_particle_config(ibm_method="DF", collision_model="BGK", delta_type="peskin4pt",
                             rotation_coupling="semi_implicit", enable_rotation=True,
                             mdf_iterations=20, time_integrator="euler_explicit",
                             marker_spacing_factor=0.83, incompressible_lbgk=True,
                             settling_inertia_model="explicit_history",
                             sedimentation_stop_at_contact=False,
                             sedimentation_stop_offset_d=2.0,
                             sedimentation_reference_basis="standard",
                             streamwise_extent_factor=1.0, particle_selection="both",
                             grid_refinement_factor=1.0):
    """The two-disc settling benchmark of ``make_two_particle_config_reference_benchmark``
    (``scenarios/sedimentation.py:489-652``): discs of d = 0.2 m at density ratios 1.5 (heavy)
    and 1.25 (light) in a 2 m wide channel, g = 9.81 m/s^2, released at (0.8, -0.13) and
    (1.2, 0.13) m and settling along +x (the channel is turned on its side).

    * dx = 0.0025 / kappa m and dt = 2.344e-4 / kappa^2 s (kappa = ``grid_refinement_factor``;
      tau = 0.59 at every kappa), so NN = round(2 / dx) + 1 (801 at kappa = 1) and the
      lattice diameter is 0.1 (NN - 1);
    * the channel is 5 widths long times ``streamwise_extent_factor``;
    * g_lat = g dt^2 / dx, the heavy disc's terminal-speed estimate
      u_t = sqrt(0.5 g_lat D) sets ``lattice_u`` and ``Re``;
    * max_steps = round(60000 f kappa^2), check_interval = round(500 kappa^2).

    ``particle_selection="light_only"`` keeps only the light disc (one-element particle list).
    The expressions keep the reference's operand order, so g_lat is the float its result
    files record."""
    from .config import SimConfig
    W_phys, d_phys, g_phys = 2.0, 0.2, 9.81
    rho_heavy, rho_light = 1.5, 1.25
    delta_rho_heavy = rho_heavy - 1.0
    if grid_refinement_factor <= 0:
        raise ValueError(f"grid_refinement_factor must be > 0 (got {grid_refinement_factor})")
    dx_phys = 0.0025 / grid_refinement_factor
    dt_phys = 2.344e-4 / grid_refinement_factor**2
    if abs(2.0 / dx_phys - round(2.0 / dx_phys)) > 1e-9:
        raise ValueError(f"grid_refinement_factor={grid_refinement_factor} does not give an "
                         "integer lattice")
    nu_lat_target = (0.59 - 0.5) / 3.0
    NN = int(round(W_phys / dx_phys)) + 1
    D_ratio = d_phys / W_phys
    lattice_D = D_ratio * (NN - 1)
    xmax = (10.0 / W_phys) * streamwise_extent_factor
    g_lattice = g_phys * dt_phys**2 / dx_phys
    u_t_est = float(np.sqrt(delta_rho_heavy * g_lattice * lattice_D))
    Re_est = u_t_est * lattice_D / nu_lat_target
    heavy = {"center": (0.8 / W_phys, (-0.13 + 1.0) / 2.0), "rho_ratio": rho_heavy}
    light = {"center": (1.2 / W_phys, (0.13 + 1.0) / 2.0), "rho_ratio": rho_light}
    if particle_selection == "both":
        particles = [heavy, light]
    elif particle_selection == "light_only":
        particles = [light]
    else:
        raise ValueError(f"particle_selection must be 'both' or 'light_only' "
                         f"(got {particle_selection!r})")
    max_steps = int(round(60000 * streamwise_extent_factor * grid_refinement_factor**2))
    check_interval = int(round(500 * grid_refinement_factor**2))
    return SimConfig(
        Re=Re_est, NN=NN, xmax=xmax, ymax=1.0, cylinder_center=particles[0]["center"],
        cylinder_D_ratio=D_ratio, lattice_u=max(u_t_est, 0.01), inflow_u=0.0,
        bc_type="settling_channel", delta_type=delta_type, ibm_method=ibm_method,
        collision_model=collision_model, mdf_iterations=mdf_iterations, max_steps=max_steps,
        use_convergence=False, check_interval=check_interval, motion_type="sedimentation",
        rho_ratio=particles[0]["rho_ratio"], gravity=g_lattice, enable_rotation=enable_rotation,
        rotation_coupling=rotation_coupling, gravity_direction="right",
        particles_config=particles, marker_spacing_factor=marker_spacing_factor,
        nx_formula="standard", time_integrator=time_integrator,
        incompressible_lbgk=incompressible_lbgk, settling_inertia_model=settling_inertia_model,
        sedimentation_stop_at_contact=sedimentation_stop_at_contact,
        sedimentation_stop_offset_d=sedimentation_stop_offset_d,
        sedimentation_reference_basis=sedimentation_reference_basis)
