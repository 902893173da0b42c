"""Ensembles of settling-particle runs stepped together on one device.

The reference runs each settling case as its own process, one particle loop per step on the
host (``iblbm/runtime/scenarios/sedimentation_multi.py:49-124``,
``sedimentation_single.py:91-293``).  Here every member is an instance of one context, so a
time step is:

1. one vectorised Verlet half step of every particle of every member (``ParticleBank``);
2. one batched marker upload for all members (``iblbm_set_markers_batch``: one copy, one
   scatter launch, one table launch, one synchronisation);
3. one step launch over all instances;
4. one batched read of the marker forces (``iblbm_download_marker_forces_batch``) for DFC and
   several-particle DF members; window-force members (single-particle DF/MDF, coupled MDF)
   are read one by one;
5. one vectorised Verlet full step.

With the explicit Euler integrator (``sedimentation_multi.py:127-202``) step 1 is empty (the
rings were placed by the previous step's update) and step 5 is one vectorised Euler update,
which reads each live disc's rho and U box (``iblbm_download_state_box``) for the fluid's
angular momentum inside it, then places every ring for the next step.

``ParticleBank`` evaluates the reference's per-particle expressions on arrays over all
particles, operation for operation in the same order, so every member's trajectory carries
the bits of its own single run (tested against ``Sedimentation`` and the reference fixtures).
"""
from __future__ import annotations

import numpy as np

from .sedimentation import (Sedimentation, compute_net_gravity, l_int_box, l_int_z,
                            stop_final_step)
from .solver import Context


def _trig_vector_matches_scalar() -> bool:
    """Whether NumPy's cos/sin over an array give, element for element, the bits of the
    same calls on single floats (the form the per-particle marker update uses)."""
    x = np.random.default_rng(7).uniform(-1.0, 1.0, 4096)
    x[:1024] *= 1e-4
    vc, vs = np.cos(x), np.sin(x)
    sc = np.array([np.cos(float(v)) for v in x])
    ss = np.array([np.sin(float(v)) for v in x])
    return bool(np.array_equal(vc.view(np.int64), sc.view(np.int64)) and
                np.array_equal(vs.view(np.int64), ss.view(np.int64)))


_VEC_TRIG = None


def _cos_sin(a: np.ndarray):
    global _VEC_TRIG
    if _VEC_TRIG is None:
        _VEC_TRIG = _trig_vector_matches_scalar()
    if _VEC_TRIG:
        return np.cos(a), np.sin(a)
    return (np.array([np.cos(float(v)) for v in a]), np.array([np.sin(float(v)) for v in a]))


def _uniform(members, name):
    vals = {getattr(m.cfg, name) for m in members}
    if len(vals) != 1:
        raise ValueError(f"an ensemble needs one {name} for every member (got {sorted(vals)})")
    return vals.pop()


class ParticleBank:
    """Rigid-body state of every particle of an ensemble, as arrays over particles.

    Particle p belongs to member ``member_of[p]``; a member's particles are consecutive.
    Marker rings are grouped by their marker count so that each group is one (P_g, Lb)
    array; ``dst`` maps a group's rows into the members' concatenated marker arrays (member
    order, particle order within a member: the layout of ``iblbm_set_markers_batch``)."""

    def __init__(self, members):
        self.model = _uniform(members, "settling_inertia_model")
        self.rotation = bool(_uniform(members, "enable_rotation"))
        self.semi_implicit = _uniform(members, "rotation_coupling") == "semi_implicit"
        gdir = _uniform(members, "gravity_direction")
        parts, owner, multi = [], [], []
        for mi, m in enumerate(members):
            for p in m.particles:
                parts.append(p)
                owner.append(mi)
                multi.append(m.multi)
        self.parts = parts
        self.P = len(parts)
        self.member_of = np.array(owner)
        self.multi = np.array(multi)
        self.methods = [m.cfg.ibm_method for m in members]
        st_of = [members[mi].st for mi in owner]
        self.dt = np.array([float(st.dt) for st in st_of])
        self.half_dt = 0.5 * self.dt
        self.dx = np.array([float(st.dx) for st in st_of])
        self.dy = np.array([float(st.dy) for st in st_of])
        self.mass = np.array([float(p.mass) for p in parts])
        self.I = np.array([float(p.I) for p in parts])
        self.larea = np.array([float(p.Larea) for p in parts])
        # constant per particle: the net gravity force and the history coefficient
        # (gravity.py:16-34; inertia/explicit_history.py:41-54)
        self.grav = np.array([compute_net_gravity(p.rho_ratio, p.lattice_r, members[mi].g, gdir)
                              for p, mi in zip(parts, owner)])
        self.hcoef = None
        if self.model == "explicit_history":
            self.hcoef = np.array([(np.pi * p.lattice_r**2) / p.mass for p in parts])
        # feng_b2's post-correction coefficient rho_f / rho_s (inertia/feng_b2.py:33-40)
        self.rinv = None
        if self.model == "feng_b2":
            self.rinv = np.array([1.0 / p.rho_ratio for p in parts])
        self.pos = np.array([p.pos for p in parts], dtype=float)
        self.vel = np.array([p.vel for p in parts], dtype=float)
        self.vel_prev = np.array([p.vel_prev for p in parts], dtype=float)
        self.force = np.array([p.force for p in parts], dtype=float)
        self.omega = np.array([float(p.omega) for p in parts])
        self.torque = np.array([float(p.torque) for p in parts])
        self.angle = np.array([float(p.angle) for p in parts])
        self.vel_half = np.zeros_like(self.vel)
        self.omega_half = np.zeros(self.P)
        self.live = np.ones(self.P, dtype=bool)
        # explicit Euler (sedimentation_multi.py:127-202): the fluid's angular momentum inside
        # each disc, read from its member's lattice after each step
        self.euler = _uniform(members, "time_integrator") == "euler_explicit"
        self.scheme = _uniform(members, "sedimentation_euler_update_scheme")
        self.l_int_prev = np.array([float(p.l_int_prev) for p in parts])
        self.r_domain = np.array([float(p.r_domain) for p in parts])
        self.inst = [getattr(members[mi], "k", mi) for mi in owner]  # the member's instance
        self.grid = [(members[mi].st.nx, members[mi].st.ny) for mi in owner]
        # marker layout
        lbs = np.array([p.Lb for p in parts])
        starts = np.concatenate([[0], np.cumsum(lbs)[:-1]])
        self.n_markers = int(lbs.sum())
        self.groups = []
        for lb in sorted(set(lbs.tolist())):
            rows = np.nonzero(lbs == lb)[0]
            self.groups.append({
                "rows": rows,
                "dst": starts[rows][:, None] + np.arange(lb)[None, :],
                "Lx_c": np.array([parts[r].Lx_c for r in rows]),
                "Ly_c": np.array([parts[r].Ly_c for r in rows]),
                "cx0": np.array([float(parts[r].cx_init) for r in rows]),
                "cy0": np.array([float(parts[r].cy_init) for r in rows]),
                "Lx": np.array([parts[r].Lx for r in rows]),
                "Ly": np.array([parts[r].Ly for r in rows]),
                "dfc": np.array([self.methods[owner[r]] == "DFC" for r in rows]),
            })
            g = self.groups[-1]
            flat = g["dst"].ravel()
            # a group whose rings are consecutive in the batch layout is one slice of it
            g["span"] = ((int(flat[0]), int(flat[-1]) + 1)
                         if np.array_equal(flat, np.arange(flat[0], flat[0] + flat.size))
                         else None)
            g["all_dfc"], g["no_dfc"] = bool(g["dfc"].all()), not bool(g["dfc"].any())
        self.Lx = np.empty(self.n_markers)
        self.Ly = np.empty(self.n_markers)
        self.ud = np.empty((self.n_markers, 2))
        if self.euler:  # the first step runs on the initial rings at rest
            for g in self.groups:
                for i, r in enumerate(g["rows"]):
                    dst = g["dst"][i]
                    self.Lx[dst] = parts[r].Lx
                    self.Ly[dst] = parts[r].Ly
                    self.ud[dst] = parts[r].desired

    def _keep_frozen(self, new, old):
        if self.live.all():
            return new
        m = self.live if new.ndim == 1 else self.live[:, None]
        return np.where(m, new, old)

    def half_step(self):
        """Verlet half step, markers and desired velocities of every particle
        (``verlet_half`` / ``update_markers`` / ``desired_velocity``, array form)."""
        F_total = self.force + self.grav
        vel_half = self.vel + (self.half_dt[:, None] * F_total) / self.mass[:, None]
        if self.hcoef is not None:
            hdv = self.hcoef[:, None] * (self.vel - self.vel_prev)
            vel_half = vel_half + 0.5 * hdv
        pos = self.pos + (self.dt[:, None] * vel_half) * self.dx[:, None]
        self.pos = self._keep_frozen(pos, self.pos)
        self.vel_half = self._keep_frozen(vel_half, self.vel_half)
        if self.rotation:  # rigid_body.py:65-74
            omega_half = self.omega + (self.half_dt * self.torque) / self.I
            self.omega_half = self._keep_frozen(omega_half, self.omega_half)
            self.angle = self._keep_frozen(self.angle + self.dt * self.omega_half, self.angle)
        else:
            self.omega_half = np.zeros(self.P)
        om = None
        if self.semi_implicit and self.rotation:
            om = np.where(self.multi, self.omega, self.omega_half)
        return self._place_markers(self.vel_half, om)

    def _place_markers(self, vel, om):
        """Every ring at its particle's position and angle, with desired velocity vel (P, 2)
        plus, where om (P,) is given and non-zero, the rotation om x r
        (``update_markers`` / ``desired_velocity``, ``markers.py:23-57``, array form)."""
        for g in self.groups:
            r = g["rows"]
            rx = g["Lx_c"] - g["cx0"][:, None]
            ry = g["Ly_c"] - g["cy0"][:, None]
            px, py = self.pos[r, 0][:, None], self.pos[r, 1][:, None]
            if self.rotation:
                c, s = _cos_sin(self.angle[r])
                c, s = c[:, None], s[:, None]
            else:  # cos(0), sin(0)
                c, s = 1.0, 0.0
            Lx = c * rx - s * ry + px
            Ly = s * rx + c * ry + py
            g["Lx"], g["Ly"] = Lx, Ly
            d0 = np.broadcast_to(vel[r, 0][:, None], Lx.shape)
            d1 = np.broadcast_to(vel[r, 1][:, None], Lx.shape)
            if om is not None:  # markers.py:41-57
                omr = om[r][:, None]
                dxp = self.dx[r][:, None]
                on = omr != 0.0
                d0 = np.where(on, d0 + (-omr * (Ly - py)) / dxp, d0)
                d1 = np.where(on, d1 + (omr * (Lx - px)) / dxp, d1)
            if g["span"] is not None:
                a, b = g["span"]
                self.Lx[a:b] = Lx.ravel()
                self.Ly[a:b] = Ly.ravel()
                ud = self.ud[a:b].reshape(Lx.shape + (2,))
                ud[:, :, 0] = d0
                ud[:, :, 1] = d1
            else:
                dst = g["dst"]
                self.Lx[dst] = Lx
                self.Ly[dst] = Ly
                self.ud[dst, 0] = d0
                self.ud[dst, 1] = d1
        return self.Lx, self.Ly, self.ud

    def forces_from_markers(self, q):
        """(force, torque) of every particle from the concatenated marker forces: DFC sums q,
        several-particle DF sums -Lf * Larea (``marker_forces_to_particles``, array form;
        each row is reduced as its own contiguous vector, NumPy's pairwise order)."""
        F = np.zeros((self.P, 2))
        T = np.zeros(self.P)
        for g in self.groups:
            r = g["rows"]
            if g["span"] is not None:
                a, b = g["span"]
                Q = q[a:b].reshape(g["dst"].shape + (2,))
            else:
                Q = q[g["dst"]]  # (P_g, Lb, 2)
            qx = np.ascontiguousarray(Q[:, :, 0])
            qy = np.ascontiguousarray(Q[:, :, 1])
            dfc = g["dfc"]
            la = self.larea[r][:, None]

            def pick(f_dfc, f_df):  # DFC sums q; several-particle DF sums -Lf * Larea
                if g["all_dfc"]:
                    return f_dfc()
                if g["no_dfc"]:
                    return f_df()
                return np.where(dfc, f_dfc(), f_df())

            F[r, 0] = pick(lambda: qx.sum(axis=1), lambda: -(qx * la).sum(axis=1))
            F[r, 1] = pick(lambda: qy.sum(axis=1), lambda: -(qy * la).sum(axis=1))
            if self.rotation:
                rx = (g["Lx"] - self.pos[r, 0][:, None]) / self.dx[r][:, None]
                ry = (g["Ly"] - self.pos[r, 1][:, None]) / self.dy[r][:, None]
                T[r] = pick(lambda: (rx * qy - ry * qx).sum(axis=1),
                            lambda: -((rx * qy - ry * qx) * la).sum(axis=1))
        return F, T

    def full_step(self, F, T):
        """Verlet full step of every particle (``verlet_full`` and ``post_correction``, array
        form)."""
        F_total = F + self.grav
        vel_new = self.vel_half + (self.half_dt[:, None] * F_total) / self.mass[:, None]
        if self.hcoef is not None:
            hdv = self.hcoef[:, None] * (self.vel - self.vel_prev)
            vel_new = vel_new + 0.5 * hdv
        if self.rotation:  # rigid_body.py:77-82
            self.omega = self._keep_frozen(self.omega_half + (self.half_dt * T) / self.I,
                                           self.omega)
            self.torque = self._keep_frozen(T, self.torque)
        if self.rinv is not None:  # post_correction
            vel_new = vel_new + self.rinv[:, None] * (self.vel - self.vel_prev)
        self.vel_prev = self._keep_frozen(self.vel.copy(), self.vel_prev)
        self.vel = self._keep_frozen(vel_new, self.vel)
        self.force = self._keep_frozen(F, self.force)

    def euler_step(self, F, T, state_box):
        """Explicit Euler step of every particle (``euler_explicit``, ``euler_rotation`` and
        ``post_correction``, array form, each element in the scalar form's order), then every
        ring and desired velocity for the next step (the new velocity, and the new rate of
        rotation whatever the coupling).  state_box(k, x0, y0, w, h) reads rho and U of
        particle k's member on a box after the step."""
        F_total = F + self.grav
        vel_new = self.vel + (self.dt[:, None] * F_total) / self.mass[:, None]
        if self.hcoef is not None:
            vel_new = vel_new + self.hcoef[:, None] * (self.vel - self.vel_prev)
        if self.scheme == "new_velocity":
            pos = self.pos + (self.dt[:, None] * vel_new) * self.dx[:, None]
        else:
            pos = self.pos + ((0.5 * self.dt)[:, None] * (self.vel + vel_new)) * self.dx[:, None]
        om = None
        if self.rotation:
            dl = np.zeros(self.P)
            for k in np.nonzero(self.live)[0]:  # l_int_z about the position before the step
                nx, ny = self.grid[k]
                cx, cy = float(self.pos[k, 0]), float(self.pos[k, 1])
                i_lo, i_hi, j_lo, j_hi = l_int_box(nx, ny, float(self.dx[k]), cx, cy,
                                                   float(self.r_domain[k]))
                ro, U = state_box(k, i_lo, j_lo, i_hi - i_lo, j_hi - j_lo)
                l_now = l_int_z(ro, U, i_lo, j_lo, float(self.dx[k]), cx, cy,
                                float(self.r_domain[k]))
                dl[k] = l_now - self.l_int_prev[k]
                self.l_int_prev[k] = l_now
            omega_new = self.omega + (dl + self.dt * T) / self.I
            if self.scheme == "new_velocity":
                inc = self.dt * omega_new
            else:
                inc = (0.5 * self.dt) * (self.omega + omega_new)
            self.angle = self._keep_frozen(self.angle + inc, self.angle)
            self.torque = self._keep_frozen(T, self.torque)
            self.omega = self._keep_frozen(omega_new, self.omega)
            om = self.omega
        if self.rinv is not None:  # post_correction
            vel_new = vel_new + self.rinv[:, None] * (self.vel - self.vel_prev)
        self.vel_prev = self._keep_frozen(self.vel.copy(), self.vel_prev)
        self.vel = self._keep_frozen(vel_new, self.vel)
        self.pos = self._keep_frozen(pos, self.pos)
        self.force = self._keep_frozen(F, self.force)
        return self._place_markers(self.vel, om)

    def to_particles(self):
        """Write the arrays back into the members' ``Particle`` objects."""
        for k, p in enumerate(self.parts):
            p.pos = self.pos[k].copy()
            p.vel = self.vel[k].copy()
            p.vel_prev = self.vel_prev[k].copy()
            p.force = self.force[k].copy()
            p.vel_half = self.vel_half[k].copy()
            p.omega = float(self.omega[k])
            p.omega_half = float(self.omega_half[k])
            p.torque = float(self.torque[k])
            p.angle = float(self.angle[k])
        for g in self.groups:
            for i, k in enumerate(g["rows"]):
                self.parts[k].Lx = g["Lx"][i].copy()
                self.parts[k].Ly = g["Ly"][i].copy()


class SedimentationEnsemble:
    """Settling-particle runs (one or several particles each, any grid, DF / MDF / DFC) as
    instances of one context, stepped together.  The members share the rigid-body options
    (inertia model, rotation, coupling, gravity direction) and the check interval.

    A member that meets a stop condition (``sedimentation_common.py:14-66``) keeps its
    particles where they are from then on and records nothing more; its
    ``termination_reason`` is kept, as the single run's would be."""

    def __init__(self, cfgs, precision="fp32dev", device=0, save_dir=None, first_index=0):
        """save_dir: each member writes the reference's status.json and
        sedimentation_history.json under save_dir/member_XXX (XXX = first_index + its
        index) at every check interval; ``finalize_outputs`` completes them."""
        if not cfgs:
            raise ValueError("an ensemble needs at least one member")
        self.ctx = Context(device)
        cache = {}
        self.members = [Sedimentation(c, device=device, precision=precision, ctx=self.ctx,
                                      f0_cache=cache) for c in cfgs]
        self.check_interval = int(_uniform(self.members, "check_interval"))
        self.bank = ParticleBank(self.members)
        self.ids = np.array([m.k for m in self.members], dtype=np.int32)
        self.lb = np.array([sum(p.Lb for p in m.particles) for m in self.members],
                           dtype=np.int32)
        self.larea = np.array([float(m.particles[0].Larea) for m in self.members])
        self.first = np.concatenate([[0], np.cumsum([len(m.particles)
                                                     for m in self.members])[:-1]])
        self.batched = all(m.cfg.ibm_method == "DFC" or (m.multi and m.cfg.ibm_method == "DF")
                           for m in self.members)
        self.ttt = 0
        self.nodes = sum(m.st.nx * m.st.ny for m in self.members)
        self.histories = [[] for _ in self.members]
        self.reasons = [None] * len(self.members)
        self.stop_steps = [None] * len(self.members)
        self.monitors = None
        if save_dir:
            import os
            from .monitor import SedimentationMonitor
            self.monitors = [SedimentationMonitor(
                os.path.join(save_dir, f"member_{first_index + mi:03d}"), m.cfg)
                for mi, m in enumerate(self.members)]
        import time
        self._t0 = time.time()

    def _forces(self):
        if self.batched:
            q = self.ctx.download_marker_forces_batch(self.ids, self.bank.n_markers)
            return self.bank.forces_from_markers(q)
        F = np.zeros((self.bank.P, 2))
        T = np.zeros(self.bank.P)
        self.bank.to_particles()
        for mi, m in enumerate(self.members):
            k0 = int(self.first[mi])
            for j, (f, t) in enumerate(m._forces()):
                F[k0 + j] = f
                T[k0 + j] = t
        return F, T

    def begin_step(self):
        """First half of a time step: particle half step, batched marker upload and the step
        launch, which runs on while the host returns."""
        self.ttt += 1
        if self.bank.euler:  # the rings the previous step's update placed (explicit Euler)
            Lx, Ly, ud = self.bank.Lx, self.bank.Ly, self.bank.ud
        else:
            Lx, Ly, ud = self.bank.half_step()
        center = self.bank.pos[self.first]
        self.ctx.set_markers_batch(self.ids, self.lb, Lx, Ly, ud, self.larea, center)
        self.ctx.step(1, diag_every=1 if self.ttt % self.check_interval == 0 else 0)

    def end_step(self):
        """Second half: the forces (waits for this context's step), the full step, the
        checks at the check interval."""
        F, T = self._forces()
        if self.bank.euler:
            bank = self.bank
            self.bank.euler_step(F, T, lambda k, x0, y0, w, h: self.ctx.download_state_box(
                bank.inst[k], x0, y0, w, h))
        else:
            self.bank.full_step(F, T)
        if self.ttt % self.check_interval == 0:
            self._check()

    def step(self):
        """One time step of every member."""
        self.begin_step()
        self.end_step()

    def _check(self):
        self.bank.to_particles()
        for mi, m in enumerate(self.members):
            if self.reasons[mi] is not None:
                continue
            m.ttt = self.ttt
            m._mdf_read = None  # the member's MDF statistics are the latest step's
            recs = m.record()
            self.histories[mi].append(recs if m.multi else recs[0])
            if self.monitors is not None:
                import time
                self.monitors[mi](self.ttt, m.particles, m.mdf_stats,
                                  elapsed=time.time() - self._t0)
            reason = m.should_stop()
            if reason is not None:
                self.reasons[mi] = reason
                self.stop_steps[mi] = self.ttt
                m.termination_reason = reason
                k0 = int(self.first[mi])
                self.bank.live[k0:k0 + len(m.particles)] = False

    def run(self, n):
        for _ in range(int(n)):
            if all(r is not None for r in self.reasons):
                break
            self.step()
        self.ctx.sync()

    def finalize_outputs(self):
        """Marks every member's status.json complete (final step, termination reason, peak
        Reynolds numbers) and writes its whole sedimentation_history.json."""
        if self.monitors is None:
            return
        for mi, (m, mon) in enumerate(zip(self.members, self.monitors)):
            mon.finalize(self.final_step(mi), self.reasons[mi] or "max_steps", tau=m.st.tau,
                         rho_ratios=[float(p.rho_ratio) for p in m.particles])

    def final_step(self, mi):
        """Member mi's ``final_step`` as the reference's loop reports it (a member stopped
        at check step s reports s - 1; one still running reports the ensemble's step)."""
        if self.reasons[mi] is None:
            return int(self.ttt)
        return stop_final_step(self.stop_steps[mi], self.reasons[mi])

    def trajectory_row(self, mi):
        """[pos x, pos y, vel x, vel y, omega, angle, force x, force y, torque] of each of
        member mi's particles, concatenated (the columns of the settling fixtures)."""
        k0 = int(self.first[mi])
        row = []
        for k in range(k0, k0 + len(self.members[mi].particles)):
            b = self.bank
            row += [b.pos[k, 0], b.pos[k, 1], b.vel[k, 0], b.vel[k, 1], b.omega[k], b.angle[k],
                    b.force[k, 0], b.force[k, 1], b.torque[k]]
        return row

    def reynolds(self):
        """Particle Reynolds number |v| D / nu of every particle (lattice units), by member."""
        out = []
        for mi, m in enumerate(self.members):
            nu = (m.st.tau - 0.5) / 3.0
            k0 = int(self.first[mi])
            v = self.bank.vel[k0:k0 + len(m.particles)]
            out.append(np.hypot(v[:, 0], v[:, 1]) * (2.0 * m.st.lattice_r) / nu)
        return out

    def close(self):
        self.ctx.close()


class PipelinedEnsembles:
    """Members split over ``parts`` contexts (each a ``SedimentationEnsemble`` on its own
    streams) whose steps are interleaved: while one part's step runs on the device, the host
    reads the forces of the part before it and moves its particles, so the host work of one
    part hides behind the device work of the others.  Members are independent, so every
    member's results are those of its single run."""

    def __init__(self, cfgs, parts=2, precision="fp32dev", device=0, save_dir=None):
        idx = np.array_split(np.arange(len(cfgs)), max(1, min(int(parts), len(cfgs))))
        self.parts = [SedimentationEnsemble([cfgs[i] for i in ix], precision=precision,
                                            device=device, save_dir=save_dir,
                                            first_index=int(ix[0]) if len(ix) else 0)
                      for ix in idx]
        self.index = [ix for ix in idx]
        self.nodes = sum(e.nodes for e in self.parts)
        self.ttt = 0
        self._open = False

    def run(self, n):
        """n time steps of every member; each part's step is launched before the previous
        part's forces are read."""
        n = int(n)
        if n <= 0:
            return
        for e in self.parts:
            e.begin_step()
        for s in range(n):
            last = s == n - 1
            for e in self.parts:
                e.end_step()
                if not last:
                    e.begin_step()
        self.ttt += n
        self.sync()

    def sync(self):
        for e in self.parts:
            e.ctx.sync()

    def autotune(self, steps=10):
        return [e.ctx.autotune(steps) for e in self.parts]

    def finalize_outputs(self):
        for e in self.parts:
            e.finalize_outputs()

    def step_launch_us(self, n=100):
        """Device time per step launch of all parts together, markers held (HIP events)."""
        for e in self.parts:
            e.ctx.sync()
        import time
        t0 = time.perf_counter()
        for _ in range(int(n)):
            for e in self.parts:
                e.ctx.step(1)
        self.sync()
        return (time.perf_counter() - t0) * 1e6 / n

    def member(self, i):
        """(part, index within the part) of member i."""
        for p, ix in enumerate(self.index):
            hit = np.nonzero(ix == i)[0]
            if hit.size:
                return self.parts[p], int(hit[0])
        raise IndexError(i)

    def close(self):
        for e in self.parts:
            e.close()


def two_disc_cfgs(n, seed=0, D=24, width_d=8, height_d=80, rho_ratio=1.01, gravity=1.5e-3,
                  Re=12.0, lattice_u=0.02, ibm_method="DFC", delta_type="hat",
                  collision_model="CM_MRT", max_steps=200_000, check_interval=200):
    """n two-disc settling cases: discs of diameter D (lattice units) and density ratio
    rho_ratio in a channel width_d D wide and height_d D tall with no-slip walls.  The
    leading disc starts on the channel axis 6 D below the top; the trailing disc sits a
    seeded vertical gap in [2D, 4D] above it with a seeded lateral offset in [0, 0.1 D]
    (numpy.random.default_rng(seed), two uniforms per case).  Re and lattice_u set the
    viscosity as in the fixed-cylinder runs (nu = lattice_u D / Re)."""
    from .config import SimConfig
    rng = np.random.default_rng(seed)
    nn = width_d * D + 1
    d_dom = 1.0 / width_d
    ymax = float(height_d) / width_d
    out = []
    for _ in range(int(n)):
        gap = rng.uniform(2.0, 4.0) * d_dom
        off = rng.uniform(0.0, 0.1) * d_dom
        lead = (0.5, ymax - 6.0 * d_dom)
        trail = (0.5 + off, lead[1] + gap)
        out.append(SimConfig(
            Re=Re, NN=nn, xmax=1.0, ymax=ymax, cylinder_D_ratio=d_dom, cylinder_center=lead,
            lattice_u=lattice_u, inflow_u=0.0, bc_type="settling_channel",
            motion_type="sedimentation", rho_ratio=rho_ratio, gravity=gravity,
            collision_model=collision_model, ibm_method=ibm_method, delta_type=delta_type,
            marker_spacing_factor=1.0, settling_inertia_model="explicit_history",
            enable_rotation=True, time_integrator="verlet",
            particles_config=[{"center": list(lead), "rho_ratio": rho_ratio},
                              {"center": list(trail), "rho_ratio": rho_ratio}],
            max_steps=max_steps, check_interval=check_interval, use_convergence=False))
    return out
