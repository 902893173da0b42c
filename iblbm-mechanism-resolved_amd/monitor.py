"""Run outputs in the reference's file formats: ``status.json`` and
``sedimentation_history.json``.

The reference writes them from its monitoring callbacks (``monitor/sedimentation.py:41-47``,
``:211-505``) and time loop (``iblbm/runtime/loop.py:109-111``), and its figure and analysis
scripts read them back (``scripts/make_fig2_cd_comparison.py:50-67``,
``scripts/make_fig7_sed_trajectory.py:72-152``).  These writers produce the same keys from
this package's runs, so those scripts read them unchanged.  Plots, frames and velocity-field
snapshots are not written.
"""
from __future__ import annotations

import json
import math
import os
from datetime import datetime


def _atomic_json(path, data, indent=2):
    tmp = path + ".tmp"
    with open(tmp, "w") as fh:
        json.dump(data, fh, indent=indent)
    os.replace(tmp, path)


def format_elapsed(seconds: float) -> str:
    """'Xh YYm ZZs' / 'Ym ZZs' / 'Zs' (``monitor/sedimentation.py:28-38``)."""
    h, m, s = int(seconds // 3600), int((seconds % 3600) // 60), int(seconds % 60)
    if h > 0:
        return f"{h}h {m:02d}m {s:02d}s"
    if m > 0:
        return f"{m}m {s:02d}s"
    return f"{s}s"


class MdfIterStats:
    """Running MDF iteration statistics of the settling status file
    (``monitor/sedimentation.py:50-107``): count, mean / max / min iterations, a histogram
    over 0..20 iterations, diverged and converged counts, mean and max final residual."""

    def __init__(self):
        self.n = 0
        self.iter_sum = 0
        self.iter_max = 0
        self.iter_min = -1
        self.diverged = 0
        self.converged = 0
        self.hist = [0] * 21
        self.res_sum = 0.0
        self.res_max = 0.0

    def add(self, stats):
        if stats is None:
            return
        ic = int(stats.get("iter_count", 0))
        res = float(stats.get("last_residual", 0.0))
        self.n += 1
        self.iter_sum += ic
        self.iter_max = max(self.iter_max, ic)
        self.iter_min = ic if self.iter_min < 0 else min(self.iter_min, ic)
        if 0 <= ic < len(self.hist):
            self.hist[ic] += 1
        self.diverged += 1 if stats.get("diverged") else 0
        self.converged += 1 if stats.get("converged") else 0
        if res != float("inf"):
            self.res_sum += res
            self.res_max = max(self.res_max, res)

    def summary(self):
        if self.n == 0:
            return None
        return {"n_steps": self.n, "iter_avg": round(self.iter_sum / self.n, 3),
                "iter_max": int(self.iter_max), "iter_min": max(int(self.iter_min), 0),
                "diverge_count": self.diverged, "converge_count": self.converged,
                "residual_avg": round(self.res_sum / self.n, 6),
                "residual_max": round(self.res_max, 6), "histogram": list(self.hist)}


def re_metrics(speed, lattice_D, nu, rho_ratio, basis="standard"):
    """Reynolds numbers of one speed (``monitor/sedimentation.py:117-127``): standard
    |v| D / nu, particle basis (times the density ratio), and the one ``basis`` names."""
    std = float(speed * lattice_D / nu) if nu > 0.0 else 0.0
    part = std * rho_ratio
    ref = part if basis == "particle_basis" else std
    return {"Re_standard": round(std, 2), "Re_particle_basis": round(part, 2),
            "Re_reference": round(ref, 2)}


def peak_re_metrics(history, lattice_D, nu, rho_ratios, basis="standard"):
    """Peak-speed Reynolds numbers per particle over a settling history and the steps they
    occur at (``monitor/sedimentation.py:130-182``)."""
    out = {"Re_standard_peaks": [], "Re_particle_basis_peaks": [], "Re_reference_peaks": [],
           "peak_steps": []}
    if not history:
        return out
    multi = "particles" in history[0]
    for pi, rr in enumerate(rho_ratios if multi else rho_ratios[:1]):
        peak, at = 0.0, 0
        for rec in history:
            p = rec["particles"][pi] if multi else rec
            v = math.sqrt(p["vx"] ** 2 + p["vy"] ** 2)
            if v > peak:
                peak, at = v, int(rec["step"])
        m = re_metrics(peak, lattice_D, nu, rr, basis)
        out["Re_standard_peaks"].append(m["Re_standard"])
        out["Re_particle_basis_peaks"].append(m["Re_particle_basis"])
        out["Re_reference_peaks"].append(m["Re_reference"])
        out["peak_steps"].append(at)
    return out


def _cfg_block_sed(cfg):
    g = lambda k, d: getattr(cfg, k, d)  # noqa: E731
    return {"rho_ratio": cfg.rho_ratio, "ibm_method": cfg.ibm_method,
            "delta_type": g("delta_type", "peskin4pt"),
            "collision_model": g("collision_model", "BGK"),
            "settling_inertia_model": g("settling_inertia_model", "unset"), "NN": cfg.NN,
            "max_steps": cfg.max_steps, "check_interval": cfg.check_interval,
            "gravity": cfg.gravity, "gravity_direction": g("gravity_direction", "down"),
            "time_integrator": g("time_integrator", "verlet"),
            "enable_rotation": g("enable_rotation", False),
            "rotation_coupling": g("rotation_coupling", "indirect"),
            "marker_spacing_factor": g("marker_spacing_factor", 1.0),
            "retraction_dx": g("retraction_dx", 0.0), "mdf_iterations": g("mdf_iterations", 1),
            "incompressible_lbgk": g("incompressible_lbgk", False),
            "sedimentation_stop_at_contact": g("sedimentation_stop_at_contact", False),
            "sedimentation_stop_offset_d": g("sedimentation_stop_offset_d", 0.0),
            "sedimentation_reference_basis": g("sedimentation_reference_basis", "standard"),
            "sedimentation_euler_update_scheme": g("sedimentation_euler_update_scheme",
                                                   "new_velocity")}


class SedimentationMonitor:
    """status.json and sedimentation_history.json of one settling run, in the format of
    ``monitor/sedimentation.py`` ``create_sedimentation_callback``.  Call it once per check
    interval with the particles; ``finalize`` marks the run complete and adds the peak
    Reynolds numbers."""

    def __init__(self, output_dir, cfg, history_every=10):
        os.makedirs(output_dir, exist_ok=True)
        self.dir, self.cfg = output_dir, cfg
        self.d_lattice = cfg.cylinder_D_ratio * (cfg.NN - 1)
        self.dx = 1.0 / (cfg.NN - 1)
        self.D_domain = self.d_lattice * self.dx
        self.x0, self.y0 = cfg.cylinder_center
        self.right = getattr(cfg, "gravity_direction", "down") == "right"
        self.multi = cfg.particles_config is not None and len(cfg.particles_config) > 1
        self.history = []
        self.mdf = MdfIterStats()
        self.calls = 0
        self.last_elapsed = 0.0
        self.every = int(history_every)

    def _u_g(self, rho_ratio):
        return math.sqrt(abs(rho_ratio - 1.0) * self.cfg.gravity * self.d_lattice)

    def _stars(self, vel, u_g):
        if u_g <= 1e-15:
            return 0.0, 0.0
        vx = float(vel[0] / u_g)
        vy = float(vel[1] / u_g) if self.right else float(-vel[1] / u_g)
        return vx, vy

    def record(self, step, particles, mdf_stats=None):
        """The history record of one check (``monitor/sedimentation.py:261-357``)."""
        if self.multi:
            pdata = []
            for pi, p in enumerate(particles):
                u_g = self._u_g(p.rho_ratio)
                vx_s, vy_s = self._stars(p.vel, u_g)
                pdata.append({"id": pi, "rho_ratio": p.rho_ratio,
                              "t_star": float(step * u_g / self.d_lattice) if u_g > 1e-15
                              else 0.0,
                              "x": float(p.pos[0]), "y": float(p.pos[1]),
                              "vx": float(p.vel[0]), "vy": float(p.vel[1]),
                              "vx_star": vx_s, "vy_star": vy_s})
            rec = {"step": int(step), "particles": pdata}
        else:
            p = particles[0]
            u_g = self._u_g(self.cfg.rho_ratio)
            vx_s, vy_s = self._stars(p.vel, u_g)
            if u_g > 1e-15:
                t_star = float(step * u_g / self.d_lattice)
                y_star = (float((p.pos[0] - self.x0) / self.D_domain) if self.right
                          else float((self.y0 - p.pos[1]) / self.D_domain))
            else:
                t_star = y_star = 0.0
            rec = {"t_star": t_star, "y_star": y_star, "vy_star": vy_s, "vx_star": vx_s,
                   "x": float(p.pos[0]), "y": float(p.pos[1]), "vx": float(p.vel[0]),
                   "vy": float(p.vel[1]), "step": int(step),
                   "f_hydro_x": float(p.force[0]), "f_hydro_y": float(p.force[1]),
                   "angle": float(p.angle), "omega": float(p.omega),
                   "torque": float(p.torque)}
        if mdf_stats is not None:
            rec["mdf_iter"] = int(mdf_stats.get("iter_count", 0))
            rec["mdf_residual"] = float(mdf_stats.get("last_residual", 0.0))
            rec["mdf_diverged"] = bool(mdf_stats.get("diverged", False))
            self.mdf.add(mdf_stats)
        return rec

    def __call__(self, step, particles, mdf_stats=None, elapsed=0.0):
        self.calls += 1
        interval = elapsed - self.last_elapsed
        self.last_elapsed = elapsed
        rec = self.record(step, particles, mdf_stats)
        self.history.append(rec)
        p0 = particles[0]
        first = rec["particles"][0] if self.multi else rec
        u_g = self._u_g(p0.rho_ratio if self.multi else self.cfg.rho_ratio)
        if self.right:
            y_star = float((p0.pos[0] - self.x0) / self.D_domain) if u_g > 1e-15 else 0.0
        else:
            y_star = float((self.y0 - p0.pos[1]) / self.D_domain) if u_g > 1e-15 else 0.0
        status = {"experiment_type": "sedimentation", "completed": False, "step": int(step),
                  "elapsed_sec": round(elapsed, 1), "elapsed_str": format_elapsed(elapsed),
                  "interval_time": round(interval, 2), "converged": False,
                  "vy_star": round(first["vy_star"], 6), "y_star": round(y_star, 4),
                  "vx_star": round(first["vx_star"], 6),
                  "particle_x": round(float(p0.pos[0]), 6),
                  "particle_y": round(float(p0.pos[1]), 6),
                  "config": _cfg_block_sed(self.cfg),
                  "updated_at": datetime.now().isoformat(timespec="seconds")}
        summ = self.mdf.summary()
        if summ is not None:
            status["mdf_iter_stats"] = summ
        _atomic_json(os.path.join(self.dir, "status.json"), status)
        if self.calls == 1 or self.calls % self.every == 0:
            self._write_history()

    def _write_history(self):
        _atomic_json(os.path.join(self.dir, "sedimentation_history.json"), self.history)

    def finalize(self, final_step=None, termination_reason=None, tau=None, rho_ratios=None,
                 converged=False):
        """Writes the whole history and marks status.json complete, with the peak Reynolds
        numbers (``monitor/sedimentation.py:434-505``)."""
        self._write_history()
        path = os.path.join(self.dir, "status.json")
        data = {}
        if os.path.exists(path):
            with open(path) as fh:
                data = json.load(fh)
        data["completed"] = True
        data["converged"] = converged
        if final_step is not None:
            data["step"] = int(final_step)
            data["final_step"] = int(final_step)
        if termination_reason is not None:
            data["termination_reason"] = termination_reason
        if tau is not None:
            nu = (float(tau) - 0.5) / 3.0
            rr = rho_ratios or [float(self.cfg.rho_ratio)]
            basis = getattr(self.cfg, "sedimentation_reference_basis", "standard")
            peaks = peak_re_metrics(self.history, self.d_lattice, nu, rr, basis)
            data.update(peaks)
            data["history_len"] = len(self.history)
            data["re_reference_basis"] = basis
            if peaks["Re_standard_peaks"]:
                data["Re_standard_peak"] = peaks["Re_standard_peaks"][0]
                data["Re_particle_basis_peak"] = peaks["Re_particle_basis_peaks"][0]
                data["Re_reference_peak"] = peaks["Re_reference_peaks"][0]
        _atomic_json(path, data)


class FixedCylinderMonitor:
    """status.json of a fixed-cylinder run with the keys the reference's result files carry
    (``data/fixed_cylinder/*/*/status.json``): the latest Cd, Cl and convergence error, the
    Cd / Cl histories and their steps, and at the end Cd_avg / Cl_avg over the last 30 % of
    the history (``avg_start_step`` names where it starts)."""

    def __init__(self, output_dir, cfg):
        os.makedirs(output_dir, exist_ok=True)
        self.dir, self.cfg = output_dir, cfg
        self.cd, self.cl, self.steps = [], [], []
        self.last_elapsed = 0.0

    def _config(self):
        c = self.cfg
        return {"Re": c.Re, "NN": c.NN, "max_steps": c.max_steps,
                "check_interval": c.check_interval,
                "convergence_start": getattr(c, "convergence_start", 0),
                "delta_type": c.delta_type, "xmax": c.xmax, "ymax": c.ymax,
                "bc_type": c.bc_type, "ibm_method": c.ibm_method,
                "collision_model": c.collision_model, "mdf_iterations": c.mdf_iterations,
                "mdf_min_iterations": c.mdf_min_iterations,
                "marker_spacing_factor": c.marker_spacing_factor}

    def _status(self, step, elapsed, error, converged):
        interval = elapsed - self.last_elapsed
        self.last_elapsed = elapsed
        return {"step": int(step), "elapsed_sec": round(elapsed, 1),
                "elapsed_str": format_elapsed(elapsed), "interval_time": round(interval, 2),
                "interval_time_str": format_elapsed(interval),
                "Cd": round(self.cd[-1], 6) if self.cd else None,
                "Cl": round(self.cl[-1], 6) if self.cl else None, "error": error,
                "converged": bool(converged), "experiment_type": "steady",
                "config": self._config(), "Cd_history": list(self.cd),
                "Cl_history": list(self.cl), "steps": list(self.steps),
                "updated_at": datetime.now().isoformat(timespec="seconds")}

    def __call__(self, step, Cd, Cl, error=1.0, converged=False, elapsed=0.0):
        self.cd.append(float(Cd))
        self.cl.append(float(Cl))
        self.steps.append(int(step))
        _atomic_json(os.path.join(self.dir, "status.json"),
                     self._status(step, elapsed, error, converged))

    def finalize(self, final_step, termination_reason, converged=False, elapsed=None):
        data = self._status(final_step, self.last_elapsed if elapsed is None else elapsed,
                            1.0, converged)
        if self.cd:
            # the last 30 % of the samples: the reference's result files start the average at
            # the first sample past 0.7 of the run (avg_start_step 70200 of 100000, 14800 of
            # 21000, 42200 of 60000 in data/fixed_cylinder/re*/*/status.json)
            k0 = int(0.7 * len(self.cd))
            data["Cd_avg"] = round(sum(self.cd[k0:]) / (len(self.cd) - k0), 6)
            data["Cl_avg"] = round(sum(self.cl[k0:]) / (len(self.cl) - k0), 6)
            data["avg_start_step"] = self.steps[k0]
        data.update({"completed": True, "final_step": int(final_step),
                     "termination_reason": termination_reason})
        _atomic_json(os.path.join(self.dir, "status.json"), data)
