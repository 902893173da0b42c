"""Run outputs in the reference's formats (status.json, sedimentation_history.json): the keys
the reference's writers produce and its figure scripts read back, and the statistics they
carry (MDF iteration summary, peak Reynolds numbers, tail averages), on synthetic runs.
CPU only."""
import json
import math
import os
from types import SimpleNamespace

import numpy as np

from conftest import golden, golden_cfg
from iblbm_amd.monitor import (FixedCylinderMonitor, MdfIterStats, SedimentationMonitor,
                               format_elapsed, peak_re_metrics)

SINGLE_KEYS = {"t_star", "y_star", "vy_star", "vx_star", "x", "y", "vx", "vy", "step",
               "f_hydro_x", "f_hydro_y"}


def _particles(n, rho, t):
    return [SimpleNamespace(pos=np.array([0.5 + 0.01 * k, 3.0 - 1e-3 * t * (k + 1)]),
                            vel=np.array([0.0, -1e-3 * (k + 1) * min(t, 50) / 50]),
                            force=np.array([1e-4, 2e-3]), angle=0.0, omega=0.0, torque=0.0,
                            rho_ratio=rho[k]) for k in range(n)]


def test_single_particle_outputs(tmp_path):
    cfg = golden_cfg(golden("sed_single_mdf_cm_p4")[1])
    mon = SedimentationMonitor(str(tmp_path), cfg, history_every=3)
    for t in range(1, 8):
        mon(t * 100, _particles(1, [cfg.rho_ratio], t),
            {"iter_count": 5 + t % 3, "last_residual": 1e-5 * t, "diverged": False,
             "converged": True}, elapsed=t * 1.5)
    st = json.load(open(tmp_path / "status.json"))
    assert st["experiment_type"] == "sedimentation" and st["completed"] is False
    assert st["step"] == 700 and st["config"]["NN"] == cfg.NN
    assert st["mdf_iter_stats"]["n_steps"] == 7
    assert st["mdf_iter_stats"]["iter_max"] == 7 and st["mdf_iter_stats"]["iter_min"] == 5
    assert sum(st["mdf_iter_stats"]["histogram"]) == 7
    mon.finalize(700, "max_steps", tau=0.6, rho_ratios=[cfg.rho_ratio])
    st = json.load(open(tmp_path / "status.json"))
    h = json.load(open(tmp_path / "sedimentation_history.json"))
    assert st["completed"] and st["final_step"] == 700 and len(h) == 7
    assert SINGLE_KEYS <= set(h[0])
    d_lat = cfg.cylinder_D_ratio * (cfg.NN - 1)
    u_g = math.sqrt((cfg.rho_ratio - 1.0) * cfg.gravity * d_lat)
    assert h[2]["vy_star"] == -h[2]["vy"] / u_g  # downward settling is positive
    peak = max(math.hypot(r["vx"], r["vy"]) for r in h)
    assert st["Re_standard_peak"] == round(peak * d_lat / ((0.6 - 0.5) / 3.0), 2)


def test_two_particle_history_reads_like_the_reference_figures(tmp_path):
    cfg = golden_cfg(golden("sed_multi_df_bgk_hat")[1])
    rho = [p["rho_ratio"] for p in cfg.particles_config]
    mon = SedimentationMonitor(str(tmp_path), cfg)
    for t in range(1, 6):
        mon(t * 10, _particles(2, rho, t))
    mon.finalize(50, "max_steps", tau=0.7, rho_ratios=rho)
    h = json.load(open(tmp_path / "sedimentation_history.json"))
    # the access pattern of scripts/make_fig7_sed_trajectory.py:72-76, :136-141
    t = np.array([s["particles"][0]["t_star"] for s in h])
    y1 = np.array([s["particles"][1]["y"] for s in h])
    v1 = np.array([np.sqrt(s["particles"][1]["vx"] ** 2 + s["particles"][1]["vy"] ** 2)
                   for s in h])
    assert np.all(np.diff(t) > 0) and np.all(np.diff(y1) < 0) and v1[-1] > 0
    st = json.load(open(tmp_path / "status.json"))
    assert len(st["Re_standard_peaks"]) == 2 and st["peak_steps"] == [50, 50]
    # particle basis = standard x density ratio (before rounding)
    assert abs(st["Re_particle_basis_peaks"][1] - st["Re_standard_peaks"][1] * rho[1]) < 0.02


def test_fixed_cylinder_status(tmp_path):
    cfg = golden_cfg(golden("run_mdf_trt_p4")[1])
    mon = FixedCylinderMonitor(str(tmp_path), cfg)
    cds = [3.0, 2.0, 1.5, 1.4, 1.3, 1.2]
    for k, cd in enumerate(cds):
        mon((k + 1) * 20, cd, 0.01 * k, elapsed=k)
    mon.finalize(120, "max_steps")
    st = json.load(open(tmp_path / "status.json"))
    assert st["completed"] and st["Cd_history"] == cds and st["steps"][-1] == 120
    assert st["Cd_avg"] == round(np.mean(cds[4:]), 6) and st["avg_start_step"] == 100
    # scripts/make_fig2_cd_comparison.py:50-51
    assert st.get("config", {}).get("check_interval", 200) == cfg.check_interval


def test_helpers():
    assert format_elapsed(3725) == "1h 02m 05s" and format_elapsed(65) == "1m 05s"
    assert format_elapsed(4.2) == "4s"
    acc = MdfIterStats()
    assert acc.summary() is None
    acc.add({"iter_count": 20, "last_residual": float("inf"), "diverged": True})
    s = acc.summary()
    assert s["diverge_count"] == 1 and s["residual_max"] == 0.0 and s["histogram"][20] == 1
    assert peak_re_metrics([], 24.0, 0.04, [1.01])["peak_steps"] == []
    assert os.path.basename(__file__) == "test_monitor.py"
