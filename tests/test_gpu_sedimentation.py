"""Settling particles on the GPU path against the reference's own runs (tests/golden/sed_*):
one particle (DF, MDF) and two (DFC spread as one ring; DF and coupled MDF spread per particle,
with stencils that share nodes); Velocity Verlet and the explicit Euler integrator (whose
rotation takes the change of the fluid's angular momentum inside the disc), compressible and
incompressible LBGK, and a disc moved at a prescribed velocity.

Trajectories (position, velocity, angular velocity, angle, force, torque per step), the MDF
statistics logged with them (iteration count, last residual, and for coupled MDF the guard trips
and omega used) and the final populations are compared bit for bit."""
import numpy as np
import pytest

from conftest import golden, golden_cfg, rel_linf
from iblbm_amd.sedimentation import Sedimentation

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["sed_single_df_bgk_hat", "sed_single_mdf_cm_p4",
                                  "sed_multi_dfc_cm_hat", "sed_multi_df_bgk_hat",
                                  "sed_multi_mdf_cm_p4", "sed_single_df_euler",
                                  "sed_multi_df_inc_verlet", "sed_multi_df_inc_euler",
                                  "sed_single_df_prescribed"])
def test_sedimentation_matches_reference(case):
    g, sc = golden(case)
    cfg = golden_cfg(sc)
    sim = Sedimentation(cfg)
    rows = []
    for _ in range(sc["final_step"]):
        sim.step()
        row = [sim.ttt]
        for p in sim.particles:
            row += [p.pos[0], p.pos[1], p.vel[0], p.vel[1], p.omega, p.angle, p.force[0],
                    p.force[1], p.torque]
        if sim.mdf_stats is not None:
            st = sim.mdf_stats
            row += [st["iter_count"], st["last_residual"]]
            if sim.multi:  # the coupled iteration of several particles
                row += [st["guard_trips"], st["omega_used"]]
        rows.append(row)
    got = np.array(rows, dtype=np.float64)
    ref = g["particles"]
    assert got.shape == ref.shape
    err = np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300), axis=0)
    print(f"{case}: max rel error per column {np.array2string(err, precision=2)}")
    # every logged column bit for bit: position, velocity, angular velocity, angle, force and
    # torque of every particle, and the MDF iteration count, residual, guard trips and omega
    bad = np.flatnonzero(np.any(got != ref, axis=0))
    assert bad.size == 0, f"columns {bad.tolist()} differ (max rel {err[bad].max():.2e})"
    f = sim.ctx.download_f(sim.k)
    print(f"{case}: populations rel Linf {rel_linf(f, g['f_post']):.3e}")
    np.testing.assert_array_equal(f, g["f_post"])


@pytest.mark.parametrize("case", ["sed_multi_mdf_guard_trip", "sed_multi_mdf_guard_maxiter",
                                  "sed_multi_mdf_guard_stop"])
def test_coupled_mdf_guard_matches_reference(case):
    """The coupled MDF of two particles started from a relaxation factor above its stable
    range (omega_use 6.4, 10.4, 32; multi_particle.py:155-165): the 1.2x residual guard trips,
    halves the factor, and either converges, runs to the iteration limit, or stops at
    omega_min (diverged).  Every logged column and the populations bit for bit."""
    g, sc = golden(case)
    sim = Sedimentation(golden_cfg(sc), mdf_omega=sc["omega_use"])
    rows = []
    for _ in range(sc["final_step"]):
        sim.step()
        row = [sim.ttt]
        for p in sim.particles:
            row += [p.pos[0], p.pos[1], p.vel[0], p.vel[1], p.omega, p.angle, p.force[0],
                    p.force[1], p.torque]
        st = sim.mdf_stats
        row += [st["iter_count"], st["last_residual"], st["guard_trips"], st["omega_used"]]
        rows.append(row)
        print(case, "step", sim.ttt, st, flush=True)
    got = np.array(rows, dtype=np.float64)
    ref = g["particles"]
    assert ref[:, -2].min() >= 1  # the guard tripped on every step
    bad = np.flatnonzero(np.any(got != ref, axis=0))
    assert bad.size == 0, f"columns {bad.tolist()} differ"
    np.testing.assert_array_equal(sim.ctx.download_f(sim.k), g["f_post"])


@pytest.mark.parametrize("case", ["sed_single_df_bgk_hat", "sed_multi_df_bgk_hat",
                                  "sed_single_mdf_cm_p4", "sed_multi_dfc_cm_hat"])
def test_state_box_matches_full_state(case):
    """``download_state_box`` (the explicit Euler integrator's read of rho and U around each
    disc) returns the sub-arrays of ``download_state``'s rho and U, the touched nodes'
    velocity-corrected U included: for the box around each disc, boxes at two corners and the
    whole grid; a box leaving the grid is refused."""
    from iblbm_amd.sedimentation import l_int_box
    g, sc = golden(case)
    sim = Sedimentation(golden_cfg(sc))
    for _ in range(5):
        sim.step()
    _, ro, U = sim.ctx.download_state(sim.k)
    nx, ny = sim.st.nx, sim.st.ny
    ro2, U2 = ro.reshape(ny, nx), U.reshape(ny, nx, 2)
    boxes = [(0, 0, nx, ny), (0, 0, 3, 2), (nx - 4, ny - 5, 4, 5)]
    for p in sim.particles:
        i_lo, i_hi, j_lo, j_hi = l_int_box(nx, ny, sim.st.dx, float(p.pos[0]),
                                           float(p.pos[1]), p.r_domain)
        boxes.append((i_lo, j_lo, i_hi - i_lo, j_hi - j_lo))
    for x0, y0, w, h in boxes:
        r, u = sim.ctx.download_state_box(sim.k, x0, y0, w, h)
        np.testing.assert_array_equal(r, ro2[y0:y0 + h, x0:x0 + w])
        np.testing.assert_array_equal(u, U2[y0:y0 + h, x0:x0 + w])
    with pytest.raises(ValueError):
        sim.ctx.download_state_box(sim.k, nx - 2, 0, 3, 1)
    sim.ctx.close()


@pytest.mark.parametrize("case", ["sed_single_df_bgk_hat", "sed_single_df_euler",
                                  "sed_single_df_prescribed"])
def test_settling_channel_under_the_lds_look_ahead(case):
    """The fused launch's row queue with the LDS look-ahead (edge rows copied from their
    clamped addresses, then the settling channel's wall closures) on settling runs whose
    window takes the fused path: every logged column and the final populations equal the
    reference's run bit for bit."""
    g, sc = golden(case)
    sim = Sedimentation(golden_cfg(sc))
    sim.ctx.set_schedule(mode="fused", step_rows=3, step_order="queue")
    assert sim.ctx.step_mode() == "fused"
    rows = []
    for _ in range(sc["final_step"]):
        sim.step()
        p = sim.particles[0]
        rows.append([sim.ttt, p.pos[0], p.pos[1], p.vel[0], p.vel[1], p.omega, p.angle,
                     p.force[0], p.force[1], p.torque])
    np.testing.assert_array_equal(np.array(rows), g["particles"])
    np.testing.assert_array_equal(sim.ctx.download_f(sim.k), g["f_post"])
    sim.ctx.close()
