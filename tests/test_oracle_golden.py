"""The CPU oracle against the vectors written from the reference (tests/golden/)."""
import numpy as np
import pytest

from conftest import BC_RUN_CASES, RUN_CASES, golden, golden_cfg, perturb_of
from oracle import oracle as O

U, US = golden("units")


def test_lattice_ops_bit_exact():
    np.testing.assert_array_equal(O.feq(U["feq_ro"], U["feq_U"], False), U["feq_c"])
    np.testing.assert_array_equal(O.feq(U["feq_ro"], U["feq_U"], True), U["feq_i"])
    for inc, tag in ((False, "c"), (True, "i")):
        r, u = O.macroscopic(U["macro_f"], inc)
        np.testing.assert_array_equal(r, U[f"macro_ro_{tag}"])
        np.testing.assert_array_equal(u, U[f"macro_U_{tag}"])
    np.testing.assert_array_equal(O.stream(U["stream_old"], U["stream_post"], 8, 6),
                                  U["stream_out"])


def test_collisions_bit_exact():
    tau = US["col_tau"]
    a = (U["col_fstar"], U["col_feq"], U["col_U"], U["col_fib"])
    np.testing.assert_array_equal(O.collide_bgk(*a, tau), U["col_bgk"])
    np.testing.assert_array_equal(O.collide_trt(*a, 1 / tau, 1 / US["col_tau_minus"]),
                                  U["col_trt"])
    np.testing.assert_array_equal(O.collide_cm(U["col_fstar"], U["col_U"], U["col_fib"],
                                               U["col_cm_sdiag"]), U["col_cm"])


def test_delta_and_boundaries_bit_exact():
    np.testing.assert_array_equal(O.delta("hat", U["delta_r"]), U["delta_hat"])
    np.testing.assert_array_equal(O.delta("peskin4pt", U["delta_r"]), U["delta_p4"])
    ny, nx = US["bc_ny"], US["bc_nx"]
    np.testing.assert_array_equal(O.bc_dirichlet_neumann(U["bc_in"], nx, ny, 0.05, 1.0),
                                  U["bc_dn"])
    np.testing.assert_array_equal(O.bc_settling(U["bc_in"], nx, ny, 1.0, False), U["bc_sc"])


@pytest.mark.parametrize("dt", ["hat", "peskin4pt"])
def test_ibm_bit_exact(dt):
    dx, La, ny, nx = US["ibm_dx"], US["ibm_Larea"], US["ibm_ny"], US["ibm_nx"]
    Lx, Ly, ud = U["ibm_Lx"], U["ibm_Ly"], U["ibm_ud"]
    for inc in (False, True):
        k = f"df_{dt}_{'i' if inc else 'c'}"
        res = O.direct_forcing(Lx, Ly, ud, U["ibm_Eux"], U["ibm_Euy"], U["ibm_Ro"], dx, dx, 1.0,
                               La, ny, nx, dt, inc)
        for nm, v in zip(["fib", "Lux", "Luy", "Lfx", "Lfy", "R"], res):
            np.testing.assert_array_equal(v, U[f"{k}_{nm}"])
    om = O.omega_from_type(Lx, Ly, dt, La, dx)
    assert om == US[f"omega_{dt}"]
    Uf = np.stack([U["ibm_Eux"].ravel(), U["ibm_Euy"].ravel()], 1)
    for n_iter, min_iter, tol in ((20, 5, 1e-4), (20, 3, 1e-2), (1, 1, 1e-4), (6, 2, 1e-12)):
        fib, st = O.multi_direct_forcing(Lx, Ly, ud, Uf, U["ibm_Ro"].ravel(), dx, dx, 1.0, La,
                                         ny, nx, n_iter, min_iter, dt, om, tol)
        k = f"mdf_{dt}_{n_iter}_{min_iter}"
        np.testing.assert_array_equal(fib, U[k + "_fib"])
        assert st == US[k + "_stats"]
    fib, st = O.multi_direct_forcing(Lx, Ly, ud, Uf, U["ibm_Ro"].ravel(), dx, dx, 1.0, La, ny,
                                     nx, 20, 5, dt, 6 * om, 1e-4)
    np.testing.assert_array_equal(fib, U[f"mdf_{dt}_div_fib"])
    assert st == US[f"mdf_{dt}_div_stats"] and st["diverged"]
    lam = O.dfc_lambda(Lx, Ly, 1.0, La, dx, dx, ny, nx, dt)
    np.testing.assert_array_equal(lam, U[f"dfc_{dt}_lambda"])
    fc, force = O.dfc_correction(Lx, Ly, ud, U[f"dfc_{dt}_fstar"], lam, dx, dx, La, ny, nx, dt)
    np.testing.assert_array_equal(fc, U[f"dfc_{dt}_fcorr"])
    np.testing.assert_array_equal(force, U[f"dfc_{dt}_force"])


@pytest.mark.parametrize("case", RUN_CASES + BC_RUN_CASES)
def test_whole_run_bit_exact(case):
    g, sc = golden(case)
    cfg = golden_cfg(sc)
    sim = O.FixedCylinder(cfg, perturb_of(sc))
    mdf = []
    for _ in range(cfg.max_steps):
        sim.step()
        if sim.mdf_stats is not None and sim.ttt % cfg.check_interval == 0:
            s = sim.mdf_stats
            mdf.append([sim.ttt, s["iter_count"], s["last_residual"], s["diverged"],
                        s["converged"]])
    st = sim.st
    for nm, a in (("fstar", st.fstar), ("f_post", sim.f_post), ("ro", st.ro), ("U", st.U),
                  ("fib", st.fib), ("Cd", sim.Cd), ("Cl", sim.Cl)):
        np.testing.assert_array_equal(np.asarray(a), g[nm], err_msg=nm)
    if "mdf_stats" in g:
        np.testing.assert_array_equal(np.array(mdf, dtype=np.float64), g["mdf_stats"])
        assert sim.omega == sc["omega"]
    if "dfc_force" in g:
        np.testing.assert_array_equal(sim.dfc_force, g["dfc_force"])
    if "lambda" in g:
        np.testing.assert_array_equal(sim.lambda_cache, g["lambda"])
