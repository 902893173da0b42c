"""The batched 72-case matrix (Re {20, 40, 100, 200} x {DF, MDF, DFC} x {hat, Peskin-4} x
{BGK, TRT, CM-MRT}) and large generic-path marker windows against the oracle.

* The matrix at 1/16 of its size in each direction (Ny 41..121, Nx = 2.5 (Ny - 1) + 1, the
  same Re, methods, kernels and models), all 72 cases in one context with their own seeded
  perturbations, under both two-kernel launch modes and after the autotuner: every member's
  populations, post-stream state, fib, MDF statistics and DFC forces bit for bit.
* Windows above the fused path's 128 markers (the generic window kernel with global working
  arrays, which most of the full-size matrix runs): 151 markers, and the matrix's largest
  window at full size (Ny = 1921, D = 120, 377 markers).
* FP32-deviation storage of the shrunk matrix against FP64 within stated bounds.
"""
import numpy as np
import pytest

from conftest import rel_linf
from iblbm_amd.ensemble import Ensemble, case_matrix, cylinder_cfg
from iblbm_amd.solver import FixedCylinder
from oracle import oracle as O

pytestmark = pytest.mark.gpu
STEPS = 40


def _check_member(sim, ref, name):
    """Every per-step output of one instance against the oracle's state, bit for bit."""
    k = sim.k
    np.testing.assert_array_equal(sim.ctx.download_f(k), ref.f_post, err_msg=name + " f")
    fs, ro, U = sim.ctx.download_state(k)
    np.testing.assert_array_equal(fs, ref.st.fstar, err_msg=name + " fstar")
    np.testing.assert_array_equal(ro, ref.st.ro, err_msg=name + " rho")
    np.testing.assert_array_equal(U, ref.st.U, err_msg=name + " U")
    if ref.method in ("DF", "MDF"):
        np.testing.assert_array_equal(sim.ctx.download_fib(k), ref.st.fib, err_msg=name + " fib")
    if ref.method == "MDF":
        ms = sim.ctx.mdf_stats(k)
        for key in ("iter_count", "last_residual", "diverged", "converged"):
            assert ms[key] == ref.mdf_stats[key], (name, key, ms[key], ref.mdf_stats[key])
    if ref.method == "DFC":
        q = sim.ctx.download_dfc_force(k, ref.st.Lb)
        np.testing.assert_array_equal(q, ref.dfc_force, err_msg=name + " q")


def _name(cfg):
    return f"Re{cfg.Re:g}/{cfg.ibm_method}/{cfg.delta_type}/{cfg.collision_model}/{cfg.NN}"


@pytest.fixture(scope="module")
def matrix_refs():
    cases = case_matrix(shrink=16)
    assert len(cases) == 72
    refs = [O.FixedCylinder(cfg, perturb=(1000 + k, 1e-3)).run(STEPS)
            for k, cfg in enumerate(cases)]
    return cases, refs


@pytest.mark.parametrize("schedule", [
    dict(mode="two_streams", bulk_threads=512, bulk_rows=1, bulk_waves=0),
    dict(mode="serial", bulk_threads=256, bulk_rows=2, bulk_waves=0, stores="nontemporal"),
    "autotune",
])
def test_matrix72_shrunk_matches_oracle(matrix_refs, schedule):
    cases, refs = matrix_refs
    print(f"72-case matrix (1/16 size), schedule {schedule}", flush=True)
    ens = Ensemble(cases, precision="fp64", init="host")
    if schedule == "autotune":
        print("autotuned:", ens.autotune(4), flush=True)
    else:
        ens.ctx.set_schedule(**schedule)
    ens.advance(STEPS, graph=True)
    for sim, ref, cfg in zip(ens.members, refs, cases):
        _check_member(sim, ref, _name(cfg))
    ens.close()


def test_window_above_fused_limit_matches_oracle():
    # 151 markers (D = 0.3 of a 161-node channel): the generic window kernel
    cfg = cylinder_cfg(100.0, 161, "MDF", "peskin4pt", "TRT")
    cfg.cylinder_D_ratio = 0.3
    steps = 100
    sim = FixedCylinder(cfg, perturb=(5, 1e-3))
    print(f"generic window: {sim.st.nx}x{sim.st.ny}, Lb={sim.st.Lb}, "
          f"launch {sim.ctx.step_mode()}", flush=True)
    assert sim.st.Lb > 128 and sim.ctx.step_mode() != "fused"
    sim.advance(steps, graph=True)
    ref = O.FixedCylinder(cfg, perturb=(5, 1e-3)).run(steps)
    _check_member(sim, ref, "Lb151 MDF/P4/TRT")
    sim.ctx.close()


@pytest.mark.parametrize("method,model,steps", [("MDF", "CM_MRT", 20), ("DFC", "TRT", 10)])
def test_full_size_largest_window_matches_oracle(method, model, steps):
    # the matrix's largest grid and window: Ny = 1921, Nx = 4801, D = 120 (377 markers)
    cfg = cylinder_cfg(200.0, 1921, method, "peskin4pt", model)
    O.set_threads(16)
    sim = FixedCylinder(cfg, perturb=(1071, 1e-3))
    print(f"full-size window: {sim.st.nx}x{sim.st.ny}, Lb={sim.st.Lb}, {method}/{model}, "
          f"{steps} steps", flush=True)
    assert (sim.st.nx, sim.st.ny) == (4801, 1921) and sim.st.Lb > 300
    sim.advance(steps, graph=True)
    ref = O.FixedCylinder(cfg, perturb=(1071, 1e-3)).run(steps)
    _check_member(sim, ref, f"D120 {method}/{model}")
    sim.ctx.close()


def test_matrix72_shrunk_fp32dev_tracks_fp64():
    """FP32-deviation storage (populations stored as f - w in FP32, window stage in FP64)
    against FP64 from the same initial states: populations within 1e-5 relative L-infinity
    and Cd within 1e-4 relative after 200 steps, for all 72 members."""
    cases = case_matrix(shrink=16)
    steps = 200
    out = {}
    for prec in ("fp64", "fp32dev"):
        ens = Ensemble(cases, precision=prec, init="host")
        ens.advance(steps, diag_every=steps, graph=True)
        out[prec] = ([ens.ctx.download_f(m.k) for m in ens.members], ens.cd_cl())
        ens.close()
    worst_f, worst_cd = 0.0, 0.0
    for k, cfg in enumerate(cases):
        ef = rel_linf(out["fp32dev"][0][k], out["fp64"][0][k])
        cd32, cd64 = out["fp32dev"][1][k][0], out["fp64"][1][k][0]
        ecd = abs(cd32 - cd64) / abs(cd64)
        worst_f, worst_cd = max(worst_f, ef), max(worst_cd, ecd)
        assert ef <= 1e-5, (_name(cfg), ef)
        assert ecd <= 1e-4, (_name(cfg), ecd)
    print(f"fp32dev vs fp64, 72 members: populations rel Linf <= {worst_f:.2e}, "
          f"Cd rel <= {worst_cd:.2e}", flush=True)
