"""Copy the reference's published benchmark results into tests/golden/published.npz.

The reference ships the outputs of its long runs as data files: the Taylor-Green spatial
convergence tables (``data/taylor_green/{df,mdf,dfc}/convergence.npz`` with Peskin-4 markers,
``data/taylor_green/{df,mdf,dfc}/hat/convergence.npz`` with the hat kernel) and the
oscillating-cylinder runs at Re = 100, KC = 5 on the full 2401 x 1601 lattice for 40000 steps
(``data/oscillating_cylinder/re100_kc5/<ibm>_<collision>_<delta>/status.json``: Cd and Cl every
50 steps, Cd_peak from step 8000; ``oscillating_re100_kc5_summary_50tail.json``: the
statistics of the last half) and the fixed-cylinder runs at Re = 20, 40, 100, 200
(``data/fixed_cylinder/re*/<ibm>_<collision>_<delta>/status.json``: Cd and Cl every 200
steps, the final step, the stop reason and the averages over the last 30 %), and the
single-disc settling runs (``data/single_particle_sedimentation*/**``: the trajectory every
100 steps, the stop and the Reynolds-number peaks) and the two-disc settling runs
(``data/two_particle_sedimentation/**``: both discs' trajectories every check interval, the
stop and the Reynolds-number peaks per disc), with the Velocity Verlet or the explicit Euler
integrator.  This script
reads them as data (``numpy.load`` without pickles, ``json``) and writes the numbers into one
fixture, since the reference tree is not present where the GPU tests run.  Nothing of the
reference is executed.

Usage:  python tests/golden/make_published.py
Output: tests/golden/published.npz
"""
from __future__ import annotations

import json
import os

import numpy as np

REF = "/root/reference/data"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "published.npz")


def main():
    out = {}
    for m in ("df", "mdf", "dfc"):
        for sub, tag in (("", "p4"), ("hat", "hat")):
            z = np.load(os.path.join(REF, "taylor_green", m, sub, "convergence.npz"))
            key = f"tg_{m}_{tag}"
            out[key + "_resolutions"] = z["resolutions"].astype(np.int64)
            if "errors_ibm" in z.files:  # df/convergence.npz also holds the pure-LBM column
                out[key + "_errors"] = z["errors_ibm"]
                out[key + "_slope"] = z["slope_ibm"]
                out["tg_lbm_p4_resolutions"] = z["resolutions"].astype(np.int64)
                out["tg_lbm_p4_errors"] = z["errors_lbm"]
                out["tg_lbm_p4_slope"] = z["slope_lbm"]
            else:
                out[key + "_errors"] = z["errors"]
                out[key + "_slope"] = z["slope"]
            for extra in ("mdf_iterations", "mdf_min_iterations", "mdf_tolerance"):
                if extra in z.files:
                    out[f"{key}_{extra}"] = z[extra]
    base = os.path.join(REF, "oscillating_cylinder", "re100_kc5")
    with open(os.path.join(base, "oscillating_re100_kc5_summary_50tail.json")) as fh:
        summary = {c["name"]: c for c in json.load(fh)["cases"]}
    names = sorted(d for d in os.listdir(base) if os.path.isdir(os.path.join(base, d)))
    out["osc_names"] = np.array(names)
    for n in names:
        with open(os.path.join(base, n, "status.json")) as fh:
            s = json.load(fh)
        assert s["step"] == s["final_step"] == s["config"]["max_steps"]
        out[f"osc_{n}_steps"] = np.array(s["steps"], dtype=np.int64)
        out[f"osc_{n}_Cd"] = np.array(s["Cd_history"], dtype=np.float64)
        out[f"osc_{n}_Cl"] = np.array(s["Cl_history"], dtype=np.float64)
        out[f"osc_{n}_Cd_peak"] = np.float64(s["Cd_peak"])
        out[f"osc_{n}_validation_start_step"] = np.int64(s["validation_start_step"])
        out[f"osc_{n}_NN"] = np.int64(s["config"]["NN"])
        out[f"osc_{n}_max_steps"] = np.int64(s["config"]["max_steps"])
        c = summary[n]
        out[f"osc_{n}_tail"] = np.array([c["tail_start_step"], c["tail_n_samples"],
                                         c["Cd_peak_50tail"], c["Cd_min_50tail"],
                                         c["Cd_amp_50tail"], c["Cd_mean_50tail"],
                                         c["Cl_amp_50tail"]], dtype=np.float64)
    # fixed cylinder, Re 20 / 40 / 100 / 200 (data/fixed_cylinder/re*/<case>/status.json)
    fc = []
    for re_dir in sorted(os.listdir(os.path.join(REF, "fixed_cylinder"))):
        base = os.path.join(REF, "fixed_cylinder", re_dir)
        for n in sorted(os.listdir(base)):
            path = os.path.join(base, n, "status.json")
            if not os.path.isfile(path):
                continue
            with open(path) as fh:
                s = json.load(fh)
            key = f"fc_{re_dir}_{n}"
            fc.append(f"{re_dir}_{n}")
            c = s["config"]
            out[key + "_steps"] = np.array(s["steps"], dtype=np.int64)
            out[key + "_Cd"] = np.array(s["Cd_history"], dtype=np.float64)
            out[key + "_Cl"] = np.array(s["Cl_history"], dtype=np.float64)
            out[key + "_avg"] = np.array([s["Cd_avg"], s["Cl_avg"], s["avg_start_step"]],
                                         dtype=np.float64)
            out[key + "_run"] = np.array([c["Re"], c["NN"], c["xmax"], c["ymax"],
                                          c["max_steps"], c["check_interval"],
                                          c["mdf_iterations"], c["mdf_min_iterations"],
                                          s["final_step"], s["elapsed_sec"]], dtype=np.float64)
            out[key + "_termination"] = np.array(str(s.get("termination_reason")))
            out[key + "_delta"] = np.array(c["delta_type"])
            out[key + "_collision"] = np.array(c["collision_model"])
            out[key + "_ibm"] = np.array(c["ibm_method"])
            assert c["bc_type"] == "dirichlet_neumann"
            assert abs(c["marker_spacing_factor"] - 2.0 / 3.0) < 1e-15
    out["fc_names"] = np.array(fc)
    # single-disc settling runs (data/single_particle_sedimentation*/**/status.json and
    # sedimentation_history.json): the ones with the Velocity Verlet integrator, the none,
    # explicit-history or feng_b2 settling inertia and indirect rotation coupling
    cols = ["step", "t_star", "y_star", "vy_star", "vx_star", "x", "y", "vx", "vy",
            "f_hydro_x", "f_hydro_y", "angle", "omega", "torque", "mdf_iter", "mdf_residual"]
    sed = []
    for top in ("single_particle_sedimentation", "single_particle_sedimentation_tall60D",
                "single_particle_sedimentation_tall80D"):
        for root, _, files in sorted(os.walk(os.path.join(REF, top))):
            if "status.json" not in files or "sedimentation_history.json" not in files:
                continue
            with open(os.path.join(root, "status.json")) as fh:
                s = json.load(fh)
            c = s.get("config", {})
            if (c.get("time_integrator") not in ("verlet", "euler_explicit")
                    or c.get("rotation_coupling") != "indirect"
                    or c.get("settling_inertia_model") not in ("none", "explicit_history",
                                                               "feng_b2")):
                continue
            rel = os.path.relpath(root, REF)
            name = rel.replace("/", "__")
            with open(os.path.join(root, "sedimentation_history.json")) as fh:
                h = json.load(fh)
            key = f"sed_{name}"
            sed.append(name)
            present = [k for k in cols if k in h[0]]
            out[key + "_columns"] = np.array(present)
            out[key + "_history"] = np.array([[float(r[k]) for k in present] for r in h])
            preset = ("tall60D" if "tall60D" in top else "tall80D" if "tall80D" in top
                      else "glowinski2001")
            out[key + "_strings"] = np.array([c["ibm_method"], c["delta_type"],
                                              c["collision_model"], c["settling_inertia_model"],
                                              c["sedimentation_reference_basis"], preset,
                                              str(s.get("termination_reason"))])
            out[key + "_numbers"] = np.array(
                [c["rho_ratio"], c["NN"], c["retraction_dx"], c["gravity"], c["max_steps"],
                 c["check_interval"], c["mdf_iterations"], s["final_step"],
                 0.01 if "nu001" in rel else np.nan], dtype=np.float64)
            out[key + "_integrator"] = np.array([c["time_integrator"],
                                                 c.get("sedimentation_euler_update_scheme",
                                                       "new_velocity")])
            for pk in ("Re_standard_peaks", "Re_particle_basis_peaks", "peak_steps"):
                out[f"{key}_{pk}"] = np.array(s.get(pk, []), dtype=np.float64)
            m = s.get("mdf_iter_stats")
            if m:
                out[key + "_mdf_iter_stats"] = np.array(
                    [m["n_steps"], m["iter_avg"], m["iter_max"], m["iter_min"],
                     m["diverge_count"], m["converge_count"], m["residual_avg"],
                     m["residual_max"]], dtype=np.float64)
    out["sed_names"] = np.array(sed)
    # two-disc settling runs (data/two_particle_sedimentation/**, Velocity Verlet): per record
    # the step and, per disc, the columns below
    pcols = ["x", "y", "vx", "vy", "t_star", "vx_star", "vy_star"]
    # (the method matrix, the 60D and 80D channels, the grid-refinement runs and the light
    # disc alone: the builder arguments follow from the run's folder and grid)
    two = []
    base = os.path.join(REF, "two_particle_sedimentation")
    runs = sorted((os.path.relpath(r, base), r) for r, _, f in os.walk(base)
                  if "status.json" in f and "sedimentation_history.json" in f)
    for rel, d in runs:
        name = os.path.basename(d)
        with open(os.path.join(d, "status.json")) as fh:
            s = json.load(fh)
        c = s["config"]
        if c.get("time_integrator") not in ("verlet", "euler_explicit") or c.get(
                "settling_inertia_model") not in ("none", "explicit_history", "feng_b2"):
            continue
        extent = 1.2 if rel.startswith("extended_60D") else (
            1.6 if rel.startswith("extended_80D") else 1.0)
        kappa = (c["NN"] - 1) / 800.0
        selection = "light_only" if rel.startswith("isolated_light") else "both"
        with open(os.path.join(d, "sedimentation_history.json")) as fh:
            h = json.load(fh)
        key = f"two_{name}"
        assert name not in two
        two.append(name)
        # one disc alone is recorded in the single-particle format
        parts = (lambda r: r["particles"]) if "particles" in h[0] else (lambda r: [r])
        out[key + "_history"] = np.array(
            [[float(r["step"])] + [float(p[k]) for p in parts(r) for k in pcols] for r in h])
        out[key + "_rho"] = np.array([float(p["rho_ratio"]) for p in c.get(
            "particles_config", [])] or ([1.5, 1.25] if "particles" in h[0] else [1.25]))
        out[key + "_strings"] = np.array([c["ibm_method"], c["delta_type"],
                                          c["collision_model"], c["settling_inertia_model"],
                                          c["rotation_coupling"], c["gravity_direction"],
                                          str(s.get("termination_reason")),
                                          c["sedimentation_reference_basis"], selection, rel])
        out[key + "_numbers"] = np.array(
            [c["rho_ratio"], c["NN"], c["gravity"], c["max_steps"], c["check_interval"],
             c["mdf_iterations"], c["marker_spacing_factor"], s["final_step"],
             c["sedimentation_stop_offset_d"], extent, kappa], dtype=np.float64)
        for pk in ("Re_standard_peaks", "Re_particle_basis_peaks", "peak_steps"):
            out[f"{key}_{pk}"] = np.array(s.get(pk, []), dtype=np.float64)
        out[key + "_integrator"] = np.array([
            c["time_integrator"], c.get("sedimentation_euler_update_scheme", "new_velocity"),
            "incompressible" if c.get("incompressible_lbgk") else "compressible"])
    out["two_names"] = np.array(two)
    out["two_columns"] = np.array(pcols)
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT}: {len(out)} arrays, {os.path.getsize(OUT)} bytes")


if __name__ == "__main__":
    main()
