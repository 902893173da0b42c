"""The reference's published long-run results, reproduced on the device at their full sizes.

* Taylor-Green spatial convergence (the reference's ``scenarios/taylor_green.py`` set-up:
  [-L, L]^2 with D = L = 10, 20, 40, 80 lattice units, u0 = 0.04, Re = 10, one convective time
  L / u0, every face held at the analytic velocity, optionally a cylinder whose markers follow
  the analytic field): the L2 error of the velocity against the analytic field
  (``iblbm/diagnostics/core.py:241-263``) at each resolution and the fitted order, against the
  reference's tables (``data/taylor_green/{df,mdf,dfc}[/hat]/convergence.npz``).  The
  pure-lattice column and the hat-kernel columns reproduce the table to 1e-11 and 2e-6
  relative; the reference's Peskin-4 IBM columns come from a code state that differs from the
  shipped one (the oracle of the shipped code lands 3e-4..4e-3 from them, the device on the
  oracle's bits), so those are held to 5e-3 and their orders to 0.02.  At D = 10 and 20 the
  device's errors equal the oracle's bit for bit.
* Oscillating cylinder, Re = 100, KC = 5, on the full 2401 x 1601 lattice for 40000 steps
  (``scenarios/oscillating.py``): Cd and Cl every 50 steps, Cd_peak from step 8000 and the
  statistics of the last half, against the reference's runs
  (``data/oscillating_cylinder/re100_kc5/*/status.json`` and its 50%-tail summary).
"""
import os

import numpy as np
import pytest

from iblbm_amd import SimConfig
from iblbm_amd import taylor_green as TG
from iblbm_amd.solver import FixedCylinder

pytestmark = pytest.mark.gpu
PUB = np.load(os.path.join(os.path.dirname(__file__), "golden", "published.npz"))


def tg_config(D, method=None, delta="peskin4pt", mdf_it=1, mdf_min=1, model="BGK"):
    """``make_tg_config`` (``scenarios/taylor_green.py:16-85``)."""
    u0, Re = 0.04, 10.0
    NN = 2 * D + 1
    T = 0.5 * (NN - 1) / u0
    steps = int(round(T))
    return SimConfig(Re=Re, NN=NN, xmax=1.0, ymax=1.0, cylinder_center=(0.5, 0.5),
                     cylinder_D_ratio=0.5, lattice_u=u0, inflow_u=u0, bc_type="velocity_inlet",
                     ibm_method=method or "DF", delta_type=delta, collision_model=model,
                     max_steps=steps, convergence_threshold=1e-10, convergence_start=steps,
                     check_interval=max(1, int(T) // 10), use_convergence=False,
                     marker_spacing_factor=1.0, nx_formula="standard",
                     scenario_type="taylor_green", tg_L=1.0, tg_u0=u0, tg_T_end=T,
                     tg_with_ibm=method is not None, mdf_iterations=mdf_it,
                     mdf_min_iterations=mdf_min)


def tg_error(cfg, precision="fp64"):
    sim = FixedCylinder(cfg, precision=precision)
    sim.advance(cfg.max_steps)
    _, _, U = sim.ctx.download_state(sim.k)
    e = TG.l2_error(U, TG.analytic_velocity(cfg, sim.st.nx, sim.st.ny, cfg.max_steps * sim.st.dt))
    sim.ctx.close()
    return e


def order(dxs, errs):
    return float(np.polyfit(np.log(dxs), np.log(errs), 1)[0])


# (column, IBM method, delta, MDF iterations, MDF minimum, error tolerance)
TG_COLUMNS = [("lbm_p4", None, "peskin4pt", 1, 1, 1e-11),
              ("df_p4", "DF", "peskin4pt", 1, 1, 5e-3),
              ("mdf_p4", "MDF", "peskin4pt", 20, 5, 5e-3),
              ("dfc_p4", "DFC", "peskin4pt", 1, 1, 5e-3),
              ("df_hat", "DF", "hat", 1, 1, 2e-6),
              ("mdf_hat", "MDF", "hat", 20, 5, 2e-6),
              ("dfc_hat", "DFC", "hat", 1, 1, 2e-6)]


@pytest.mark.parametrize("col", TG_COLUMNS, ids=[c[0] for c in TG_COLUMNS])
def test_taylor_green_convergence_table(col):
    name, method, delta, mi, mn, tol = col
    res = PUB[f"tg_{name}_resolutions"]
    ref = PUB[f"tg_{name}_errors"]
    errs = np.array([tg_error(tg_config(int(D), method, delta, mi, mn)) for D in res])
    rel = np.abs(errs / ref - 1.0)
    dxs = 1.0 / (2 * res)
    print(f"TG {name}: errors {errs.tolist()} rel to the reference's table {rel.tolist()}; "
          f"order {order(dxs, errs):.4f} (table {float(PUB[f'tg_{name}_slope']):.4f})")
    assert np.all(rel <= tol)
    slope_tol = 1e-9 if tol < 1e-10 else (1e-5 if tol < 1e-5 else 0.02)
    assert abs(order(dxs, errs) - float(PUB[f"tg_{name}_slope"])) <= slope_tol
    assert order(dxs, errs) > 1.75  # second order within the reference's own scatter


@pytest.mark.parametrize("col", [TG_COLUMNS[0], TG_COLUMNS[1], TG_COLUMNS[5]],
                         ids=["lbm_p4", "df_p4", "mdf_hat"])
def test_taylor_green_matches_oracle_bitwise(col):
    import oracle.oracle as O
    name, method, delta, mi, mn, _ = col
    for D in (10, 20):
        cfg = tg_config(D, method, delta, mi, mn)
        ref = O.FixedCylinder(cfg)
        for _ in range(cfg.max_steps):
            ref.step()
        e_ref = TG.l2_error(ref.st.U, TG.analytic_velocity(cfg, ref.st.nx, ref.st.ny,
                                                           cfg.max_steps * ref.st.dt))
        assert tg_error(cfg) == e_ref


@pytest.mark.parametrize("col", [TG_COLUMNS[0], TG_COLUMNS[1], TG_COLUMNS[6]],
                         ids=["lbm_p4", "df_p4", "dfc_hat"])
def test_taylor_green_fp32dev_keeps_the_order(col):
    """FP32-deviation storage: the errors within 1% of FP64's up to D = 40, the order kept."""
    name, method, delta, mi, mn, _ = col
    res = PUB[f"tg_{name}_resolutions"][:3]
    e64 = np.array([tg_error(tg_config(int(D), method, delta, mi, mn)) for D in res])
    e32 = np.array([tg_error(tg_config(int(D), method, delta, mi, mn), "fp32dev") for D in res])
    print(f"TG {name} fp32dev: {e32.tolist()} vs fp64 {e64.tolist()}")
    assert np.all(np.abs(e32 / e64 - 1.0) <= 1e-2)
    dxs = 1.0 / (2 * res)
    assert abs(order(dxs, e32) - order(dxs, e64)) <= 0.02


def osc_config(name):
    """``make_oscillating_cylinder_config`` (``scenarios/oscillating.py:12-49``)."""
    ibm, rest = name.split("_", 1)
    delta = "hat" if rest.endswith("hat") else "peskin4pt"
    model = {"bgk": "BGK", "trt": "TRT", "cm_mrt": "CM_MRT"}[rest.rsplit("_", 1)[0]]
    return SimConfig(Re=100.0, NN=int(PUB[f"osc_{name}_NN"]), xmax=1.5, ymax=1.0,
                     cylinder_center=(0.75, 0.5), cylinder_D_ratio=1.0 / 20.0, lattice_u=0.1,
                     inflow_u=0.0, bc_type="open_boundary", delta_type=delta,
                     ibm_method=ibm.upper(), collision_model=model,
                     mdf_iterations=20 if ibm == "mdf" else 1,
                     mdf_min_iterations=5 if ibm == "mdf" else 1,
                     max_steps=int(PUB[f"osc_{name}_max_steps"]), check_interval=50,
                     use_convergence=False, motion_type="oscillating", KC=5.0,
                     marker_spacing_factor=2.0 / 3.0, nx_formula="simple")


def tail_stats(steps, cd, cl, start):
    m = steps >= start
    c, l = cd[m], cl[m]
    return np.array([m.sum(), c.max(), c.min(), 0.5 * (c.max() - c.min()), c.mean(),
                     0.5 * (l.max() - l.min())])


# one run per (IBM method, collision model, kernel) pairing that covers each once; every
# one of the reference's 18 runs with IBLBM_ALL_PUBLISHED=1 (about 12 s each)
OSC_DEFAULT = ["df_bgk_hat", "df_cm_mrt_p4", "mdf_trt_p4", "mdf_bgk_hat", "dfc_cm_mrt_hat",
               "dfc_trt_p4"]
OSC_CASES = ([str(n) for n in PUB["osc_names"]] if os.environ.get("IBLBM_ALL_PUBLISHED")
             else OSC_DEFAULT)


@pytest.mark.parametrize("name", OSC_CASES)
def test_oscillating_cylinder_full_size(name):
    cfg = osc_config(name)
    sim = FixedCylinder(cfg)
    print(f"oscillating {name}: {sim.st.nx} x {sim.st.ny}, {sim.st.Lx.size} markers, "
          f"{cfg.max_steps} steps", flush=True)
    steps, cd, cl = [], [], []
    while sim.ttt < cfg.max_steps:
        sim.advance(cfg.check_interval, diag_every=cfg.check_interval)
        c, l = sim.cd_cl()
        steps.append(sim.ttt)
        cd.append(c)
        cl.append(l)
        if sim.ttt % 5000 == 0:
            print(f"  step {sim.ttt}: Cd {c:.6f} (reference "
                  f"{PUB[f'osc_{name}_Cd'][len(cd) - 1]:.6f})", flush=True)
    steps, cd, cl = np.array(steps), np.array(cd), np.array(cl)
    sim.ctx.close()
    np.testing.assert_array_equal(steps, PUB[f"osc_{name}_steps"])
    ref_cd, ref_cl = PUB[f"osc_{name}_Cd"], PUB[f"osc_{name}_Cl"]
    dcd = np.abs(cd - ref_cd)
    peak = cd[steps >= PUB[f"osc_{name}_validation_start_step"]].max()
    tail = tail_stats(steps, cd, cl, PUB[f"osc_{name}_tail"][0])
    print(f"  max |dCd| {dcd.max():.3e} (at step {steps[dcd.argmax()]}), max |dCl| "
          f"{np.abs(cl - ref_cl).max():.3e}; Cd_peak {peak:.6f} (reference "
          f"{float(PUB[f'osc_{name}_Cd_peak']):.6f}); tail {tail.tolist()} (reference "
          f"{PUB[f'osc_{name}_tail'][1:].tolist()})")
    assert np.all(np.isfinite(cd))
    # the reference's histories are reproduced to round-off (the device sums rho and the
    # marker forces in a fixed tree order, not NumPy's); Cd_peak is published to 6 decimals
    assert dcd.max() <= 1e-9 * np.abs(ref_cd).max()
    assert np.abs(cl - ref_cl).max() <= 1e-9 * np.abs(ref_cd).max()
    assert abs(peak - float(PUB[f"osc_{name}_Cd_peak"])) <= 5.01e-7
    ref_tail = PUB[f"osc_{name}_tail"]
    assert tail[0] == ref_tail[1]
    np.testing.assert_allclose(tail[1:], ref_tail[2:], atol=1.5e-4)  # the summary's 4 decimals


def fc_config(name):
    """``make_dirichlet_neumann_fixed_cylinder_config`` (``scenarios/steady.py:14-78``) with
    the published run's own settings.  Runs whose file names a stop reason were made with the
    Cd-window stop (range of the last 5 samples below 1e-4 after step 20000); the others
    (no stop reason in the file) without it, and ran their full length."""
    key = f"fc_{name}"
    Re, NN, xmax, ymax, max_steps, ci, mi, mn, _, _ = PUB[key + "_run"]
    window_stop = str(PUB[key + "_termination"]) != "None"
    return SimConfig(Re=float(Re), NN=int(NN), xmax=float(xmax), ymax=float(ymax),
                     cylinder_center=(0.5, 0.5), cylinder_D_ratio=1.0 / 40.0, lattice_u=0.1,
                     inflow_u=0.1, bc_type="dirichlet_neumann",
                     delta_type=str(PUB[key + "_delta"]), ibm_method=str(PUB[key + "_ibm"]),
                     collision_model=str(PUB[key + "_collision"]), mdf_iterations=int(mi),
                     mdf_min_iterations=int(mn), max_steps=int(max_steps),
                     check_interval=int(ci), use_convergence=False,
                     cd_convergence_tol=1e-4 if window_stop else None,
                     cd_convergence_window=5, cd_convergence_start_step=20_000,
                     marker_spacing_factor=2.0 / 3.0, nx_formula="standard")


# the Cd-window stop (Re 20 DF BGK hat, 21000 steps), an MDF run, a DFC run and the largest
# lattice (Re 200, 3001 x 2401, 60000 steps); all 72 with IBLBM_ALL_PUBLISHED=1
FC_DEFAULT = ["re20_df_bgk_hat", "re40_mdf_cm_mrt_hat", "re100_dfc_trt_p4", "re200_df_bgk_p4"]
FC_CASES = ([str(n) for n in PUB["fc_names"]] if os.environ.get("IBLBM_ALL_PUBLISHED")
            else FC_DEFAULT)


@pytest.mark.parametrize("name", FC_CASES)
def test_fixed_cylinder_full_size(name, tmp_path):
    """``run()`` with its status writer against the reference's result file of the same run:
    the Cd / Cl histories, the final step, the stop reason and the 30 %-tail averages."""
    import json
    from iblbm_amd.solver import run
    key = f"fc_{name}"
    cfg = fc_config(name)
    print(f"fixed cylinder {name}: NN {cfg.NN}, xmax {cfg.xmax}, {cfg.ibm_method} "
          f"{cfg.collision_model} {cfg.delta_type}, up to {cfg.max_steps} steps", flush=True)
    res = run(cfg, verbose=False, save_dir=str(tmp_path))
    st = json.load(open(tmp_path / "status.json"))
    steps = np.array(st["steps"])
    cd, cl = np.array(st["Cd_history"]), np.array(st["Cl_history"])
    ref_cd, ref_cl = PUB[key + "_Cd"], PUB[key + "_Cl"]
    run_pub = PUB[key + "_run"]
    n = min(len(cd), len(ref_cd))
    dcd = np.abs(cd[:n] - ref_cd[:n])
    print(f"  {res['sim'].st.nx} x {res['sim'].st.ny}: final step {st['final_step']} "
          f"(reference {int(run_pub[8])}), stop {st['termination_reason']} (reference "
          f"{PUB[key + '_termination']}); max |dCd| {dcd.max():.3e}, max |dCl| "
          f"{np.abs(cl[:n] - ref_cl[:n]).max():.3e}; Cd_avg {st['Cd_avg']} (reference "
          f"{PUB[key + '_avg'][0]})", flush=True)
    res["sim"].ctx.close()
    np.testing.assert_array_equal(steps, PUB[key + "_steps"])
    assert st["final_step"] == int(run_pub[8])
    if str(PUB[key + "_termination"]) != "None":
        assert st["termination_reason"] == str(PUB[key + "_termination"])
    # steady wakes (Re 20, 40) keep the reference's digits to round-off over 100000 steps;
    # the shedding wakes (Re 100, 200) amplify the last-bit differences of the reference's
    # own run (made on its GPU path) to at most 5e-9 by the end of the run
    tol = (1e-12 if cfg.Re <= 40 else 1e-8) * abs(PUB[key + "_avg"][0])
    assert dcd.max() <= tol
    assert np.abs(cl[:n] - ref_cl[:n]).max() <= tol
    assert st["avg_start_step"] == int(PUB[key + "_avg"][2])
    assert abs(st["Cd_avg"] - PUB[key + "_avg"][0]) <= 1.01e-6  # 6 decimals published
    assert abs(st["Cl_avg"] - PUB[key + "_avg"][1]) <= 1.01e-6


def sed_config(name):
    """The single-disc settling configuration of the reference's runner
    (``scripts/experiments/run_single_particle_sedimentation_experiments.py:287-305`` over
    ``scenarios/sedimentation.py:36-249``) with the published run's settings."""
    from iblbm_amd.sedimentation import make_sedimentation_config
    key = f"sed_{name}"
    ibm, delta, coll, inertia, basis, preset, _ = [str(v) for v in PUB[key + "_strings"]]
    rho, NN, ret, _, _, _, mi, _, nu = PUB[key + "_numbers"]
    integrator, scheme = [str(v) for v in PUB[key + "_integrator"]]
    cfg = make_sedimentation_config(
        rho_ratio=float(rho), ibm_method=ibm, delta_type=delta, collision_model=coll,
        NN=int(NN), case_name=preset, retraction_dx=float(ret),
        nu_phys_override=None if np.isnan(nu) else float(nu), enable_rotation=True,
        rotation_coupling="indirect", mdf_iterations=int(mi), time_integrator=integrator,
        settling_inertia_model=inertia, sedimentation_stop_at_contact=True,
        sedimentation_reference_basis=basis)
    cfg.sedimentation_euler_update_scheme = scheme
    return cfg


# the MDF + Peskin-4 run (about 70 s on the device: the particle update runs on the host
# between steps); every published run within the device path with IBLBM_ALL_PUBLISHED=1
SED_DEFAULT = [
    "single_particle_sedimentation__method_matrix__rho150__mdf_bgk_verlet_explicit_history_p4",
]
SED_CASES = ([str(n) for n in PUB["sed_names"]] if os.environ.get("IBLBM_ALL_PUBLISHED")
             else SED_DEFAULT)


@pytest.mark.parametrize("name", SED_CASES)
def test_single_disc_settling_full_size(name, tmp_path):
    """``sedimentation.run`` with its status writer against the reference's result files of
    the same run: the trajectory every 100 steps (position, velocity, angle, rotation rate,
    hydrodynamic force and torque, MDF sub-steps and residual), the stop at the wall and
    the Reynolds-number peaks."""
    import json
    from iblbm_amd.sedimentation import run
    key = f"sed_{name}"
    cfg = sed_config(name)
    nums = PUB[key + "_numbers"]
    assert cfg.gravity == nums[3] and cfg.max_steps == int(nums[4])
    print(f"settling {name}: NN {cfg.NN}, ymax {cfg.ymax}, {cfg.ibm_method} "
          f"{cfg.collision_model} {cfg.delta_type} {cfg.settling_inertia_model}", flush=True)
    res = run(cfg, verbose=False, save_dir=str(tmp_path))
    h = json.load(open(tmp_path / "sedimentation_history.json"))
    st = json.load(open(tmp_path / "status.json"))
    res["sim"].ctx.close()
    cols = [str(c) for c in PUB[key + "_columns"]]
    ref = PUB[key + "_history"]
    got = np.array([[float(r[c]) for c in cols] for r in h])
    n = min(len(got), len(ref))
    worst = {}
    for k, c in enumerate(cols):
        scale = max(np.abs(ref[:n, k]).max(), 1e-300)
        worst[c] = float(np.abs(got[:n, k] - ref[:n, k]).max() / scale)
    print(f"  {len(got)} records (reference {len(ref)}), final step {st['final_step']} "
          f"(reference {int(nums[7])}), stop {st['termination_reason']}; worst column "
          f"difference / column scale: " + ", ".join(f"{c} {v:.1e}" for c, v in worst.items()),
          flush=True)
    assert len(got) == len(ref)
    np.testing.assert_array_equal(got[:, 0], ref[:, 0])  # steps
    assert st["termination_reason"] == str(PUB[key + "_strings"][6])
    # the current loop reports the step counter minus one for a run stopped by its check
    # (iblbm/runtime/loop.py:106-109, :142-148); some of the result files come from a loop
    # that reported the counter itself, on the same trajectory and check step
    ref_final = int(nums[7])
    assert st["final_step"] == ref_final or (st["final_step"] == ref_final - 1 and
                                             ref_final == int(ref[-1, 0]))
    for c, v in worst.items():
        assert v <= 1e-6, c


def two_config(name):
    """The two-disc settling configuration of the reference's runner
    (``scripts/experiments/run_two_particle_sedimentation_experiments.py:287-304`` over
    ``scenarios/sedimentation.py:489-652``): semi-implicit rotation coupling, marker spacing
    0.83, the offset stop at 2 diameters, 20 MDF sub-steps for MDF."""
    from iblbm_amd.sedimentation import make_two_particle_config
    key = f"two_{name}"
    ibm, delta, coll, inertia = [str(v) for v in PUB[key + "_strings"][:4]]
    selection = str(PUB[key + "_strings"][8])
    extent, kappa = [float(v) for v in PUB[key + "_numbers"][9:11]]
    integrator, scheme, comp = [str(v) for v in PUB[key + "_integrator"]]
    cfg = make_two_particle_config(
        ibm_method=ibm, collision_model=coll, delta_type=delta,
        rotation_coupling="semi_implicit", enable_rotation=True,
        mdf_iterations=20 if ibm == "MDF" else 1, time_integrator=integrator,
        marker_spacing_factor=0.83, incompressible_lbgk=comp == "incompressible",
        settling_inertia_model=inertia, sedimentation_stop_offset_d=2.0,
        streamwise_extent_factor=extent, particle_selection=selection,
        grid_refinement_factor=kappa)
    cfg.sedimentation_euler_update_scheme = scheme
    for k, v in TWO_SHORTENED.get(name, {}).items():
        setattr(cfg, k, v)
    return cfg


# a run its runner shortened after building the configuration (the same builder, the first
# 20,000 steps with a record every 100)
TWO_SHORTENED = {"highlight_run": dict(max_steps=20_000, check_interval=100)}


# the DF run with the feng_b2 correction (about a minute on the device) and the shortened
# explicit-Euler run (incompressible LBGK, the discs' rotation fed by the fluid's angular
# momentum inside them); every published run with IBLBM_ALL_PUBLISHED=1 (the coupled-MDF runs
# take several minutes each)
TWO_DEFAULT = ["df_bgk_verlet_feng_b2", "highlight_run"]
# A result file that its recorded settings do not reproduce: the coupled-MDF TRT run differs
# from its first record on (2.3e-3 of the column scale at step 500, where every other run of
# the family agrees to 1e-12).  Runs that take 20 MDF sub-steps every step come closest
# (1.9e-5 at step 500; 20 sub-steps with other relaxation factors, or 40 and 100 sub-steps,
# come no closer), and the reference took 2.7x as long for it as for the BGK runs: the file
# was made with more MDF work per step than its recorded settings give.
TWO_UNMATCHED = {
    "mdf_trt_verlet_explicit_history": "result file not made with its recorded settings "
                                       "(differs from step 500 on; 20 MDF sub-steps every step "
                                       "come closest)"}
# Result files whose arithmetic differs from the shipped code's at round-off from the first
# record on: the 80D channel's file differs from the reference's own 50D run of the same case
# by up to 4e-8 at steps 500-1500, before anything can tell the channels apart (its 60D file
# equals the 50D one there bit for bit); the refined-grid and light-disc files of the same
# series (they report the stop step as the 80D file does) start 1e-12..1e-10 from the device
# at the first record likewise, where the runs that match agree to 1e-13.  The differences
# grow along the run (measured at most 2e-4 of a column's scale); the stop step is the same.
# The two full explicit-Euler runs start their drift later, from the fluid's angular momentum
# inside each disc: their files come from the reference's CuPy path (about 220 MLUPS by their
# elapsed time), whose compute_l_int_z sums on the device in another order than NumPy's
# pairwise sum, and the step's change of that sum (a difference of nearly equal numbers)
# carries its last-bit differences into the rotation.  The device agrees with the NumPy path
# bit for bit on every column (tests/test_gpu_sedimentation.py, sed_multi_df_inc_euler); it
# stays within 1e-12 of these files to step 8,000, the chaotic two-disc interaction then
# grows the difference (9e-9 at step 20,000, the shortened run, 5e-6 at the end) with the same
# stop step.
TWO_DRIFT = {"df_bgk_verlet_none_80D", "df_bgk_verlet_explicit_history_ny961",
             "df_bgk_verlet_none_ny961", "df_bgk_verlet_explicit_history_ny1281",
             "df_bgk_verlet_none_ny1281", "df_bgk_verlet_light_explicit_history",
             "df_bgk_verlet_light_none", "baseline", "reference_none"}
TWO_CASES = [pytest.param(n, marks=pytest.mark.skip(reason=TWO_UNMATCHED[n]))
             if n in TWO_UNMATCHED else n
             for n in ([str(n) for n in PUB["two_names"]] if os.environ.get("IBLBM_ALL_PUBLISHED")
                       else TWO_DEFAULT)]


@pytest.mark.parametrize("name", TWO_CASES)
def test_two_disc_settling_full_size(name, tmp_path):
    """Two discs (density ratios 1.5 and 1.25, D = 80 lattice units, 303 markers each)
    settling along a 4001 x 801 channel (``sedimentation.run`` with its status writer)
    against the reference's result files of the same run: both discs' positions, velocities
    and scaled times every 500 steps, the offset stop and the Reynolds-number peaks.  The
    other published runs of the same builder: 60D and 80D channels (4801 and 6401 long),
    grids refined by 1.2 and 1.6 (961 and 1281 wide, time step scaled by the square), and the
    light disc alone."""
    import json
    from iblbm_amd.sedimentation import run
    key = f"two_{name}"
    cfg = two_config(name)
    nums = PUB[key + "_numbers"]
    assert cfg.gravity == nums[2] and cfg.max_steps == int(nums[3])
    assert [p["rho_ratio"] for p in cfg.particles_config] == list(PUB[key + "_rho"])
    print(f"two discs {name}: {cfg.ibm_method} {cfg.collision_model} {cfg.delta_type} "
          f"{cfg.settling_inertia_model}", flush=True)
    res = run(cfg, verbose=True, save_dir=str(tmp_path))  # a line per record: long runs
    h = json.load(open(tmp_path / "sedimentation_history.json"))
    st = json.load(open(tmp_path / "status.json"))
    res["sim"].ctx.close()
    cols = [str(c) for c in PUB["two_columns"]]
    ref = PUB[key + "_history"]
    # one disc alone is recorded in the single-particle format
    got = np.array([[float(r["step"])] + [float(p[c]) for p in r.get("particles", [r])
                                          for c in cols] for r in h])
    n = min(len(got), len(ref))
    worst = np.abs(got[:n, 1:] - ref[:n, 1:]).max(axis=0) / np.maximum(
        np.abs(ref[:n, 1:]).max(axis=0), 1e-300)
    labels = [f"{c}{i}" for i in range(len(cfg.particles_config)) for c in cols]
    per_rec = (np.abs(got[:n, 1:] - ref[:n, 1:]) /
               np.maximum(np.abs(ref[:n, 1:]).max(axis=0), 1e-300)).max(axis=1)
    grow = {f"{e:.0e}": (int(got[np.argmax(per_rec > e), 0]) if (per_rec > e).any() else None)
            for e in (1e-12, 1e-10, 1e-8, 1e-6)}
    quarters = [float(f"{per_rec[int(q * (n - 1))]:.1e}") for q in (0.25, 0.5, 0.75, 1.0)]
    print(f"  first step where the worst column difference exceeds: {grow}; at the quarters "
          f"of the run: {quarters}", flush=True)
    print(f"  {len(got)} records (reference {len(ref)}), final step {st['final_step']} "
          f"(reference {int(nums[7])}), stop {st['termination_reason']}; worst column "
          f"difference / column scale: " + ", ".join(f"{c} {v:.1e}" for c, v in
                                                      zip(labels, worst)), flush=True)
    assert len(got) == len(ref)
    np.testing.assert_array_equal(got[:, 0], ref[:, 0])  # steps
    assert st["termination_reason"] == str(PUB[key + "_strings"][6])
    # the current loop reports the step counter minus one for a run stopped by its check
    # (iblbm/runtime/loop.py:106-109, :142-148); some of the result files come from a loop
    # that reported the counter itself, on the same trajectory and check step
    ref_final = int(nums[7])
    assert st["final_step"] == ref_final or (st["final_step"] == ref_final - 1 and
                                             ref_final == int(ref[-1, 0]))
    drift = name in TWO_DRIFT
    assert worst.max() <= (1e-3 if drift else 1e-6), labels[int(np.argmax(worst))]
    # peak Reynolds numbers per disc, published to two decimals
    for pk in ("Re_standard_peaks", "Re_particle_basis_peaks"):
        if drift:
            np.testing.assert_allclose(st[pk], PUB[f"{key}_{pk}"], rtol=2e-3, atol=0)
        else:
            np.testing.assert_allclose(st[pk], PUB[f"{key}_{pk}"], atol=0.0051, rtol=0)
    if not drift:
        np.testing.assert_array_equal(st["peak_steps"], PUB[key + "_peak_steps"])
