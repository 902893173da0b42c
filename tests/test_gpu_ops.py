"""Operator-level HIP kernels against the reference vectors: bit-exact FP64."""
import numpy as np
import pytest

from conftest import golden
from iblbm_amd import gpu_kernels as G

pytestmark = pytest.mark.gpu
U, US = golden("units")


def test_feq_macro_stream():
    np.testing.assert_array_equal(G.feq_gpu(U["feq_ro"], U["feq_U"], False), U["feq_c"])
    np.testing.assert_array_equal(G.feq_gpu(U["feq_ro"], U["feq_U"], True), U["feq_i"])
    for inc, tag in ((False, "c"), (True, "i")):
        r, u = G.macroscopic_gpu(U["macro_f"], inc)
        np.testing.assert_array_equal(r, U[f"macro_ro_{tag}"])
        np.testing.assert_array_equal(u, U[f"macro_U_{tag}"])
    np.testing.assert_array_equal(G.streaming_gpu(U["stream_old"], U["stream_post"], 8, 6),
                                  U["stream_out"])


def test_collisions():
    tau = US["col_tau"]
    a = (U["col_fstar"], U["col_feq"], U["col_U"], U["col_fib"])
    np.testing.assert_array_equal(G.collision_gpu(*a, tau, 1.0), U["col_bgk"])
    np.testing.assert_array_equal(G.collision_trt_gpu(*a, tau, US["col_tau_minus"], 1.0),
                                  U["col_trt"])
    np.testing.assert_array_equal(G.collision_cm_mrt_gpu(*a, U["col_cm_sdiag"], 1.0),
                                  U["col_cm"])


def test_boundaries():
    ny, nx = US["bc_ny"], US["bc_nx"]
    np.testing.assert_array_equal(G.boundary_gpu("dirichlet_neumann", U["bc_in"], nx, ny, 0.05),
                                  U["bc_dn"])
    np.testing.assert_array_equal(G.boundary_gpu("settling_channel", U["bc_in"], nx, ny, 0.0),
                                  U["bc_sc"])


@pytest.mark.parametrize("dt", ["hat", "peskin4pt"])
def test_direct_forcing(dt):
    dx, La, ny, nx = US["ibm_dx"], US["ibm_Larea"], US["ibm_ny"], US["ibm_nx"]
    for inc in (False, True):
        k = f"df_{dt}_{'i' if inc else 'c'}"
        fib, Lux, Luy = G.ibm_direct_forcing_gpu(U["ibm_Lx"], U["ibm_Ly"], U["ibm_ud"],
                                                 U["ibm_Eux"], U["ibm_Euy"], U["ibm_Ro"], dx,
                                                 dx, 1.0, La, ny, nx, dt, True, inc)
        np.testing.assert_array_equal(fib, U[k + "_fib"])
        np.testing.assert_array_equal(Lux, U[k + "_Lux"])
        np.testing.assert_array_equal(Luy, U[k + "_Luy"])


@pytest.mark.parametrize("dt", ["hat", "peskin4pt"])
def test_multi_direct_forcing(dt):
    dx, La, ny, nx = US["ibm_dx"], US["ibm_Larea"], US["ibm_ny"], US["ibm_nx"]
    om = US[f"omega_{dt}"]
    Uf = np.stack([U["ibm_Eux"].ravel(), U["ibm_Euy"].ravel()], 1)
    runs = [((20, 5, 1e-4), f"mdf_{dt}_20_5"), ((20, 3, 1e-2), f"mdf_{dt}_20_3"),
            ((1, 1, 1e-4), f"mdf_{dt}_1_1"), ((6, 2, 1e-12), f"mdf_{dt}_6_2")]
    for (n_iter, min_iter, tol), k in runs:
        fib, st = G.ibm_multi_direct_forcing_gpu(U["ibm_Lx"], U["ibm_Ly"], U["ibm_ud"], Uf,
                                                 U["ibm_Ro"].ravel(), dx, dx, 1.0, La, ny, nx,
                                                 n_iter, min_iter, dt, om, tol)
        np.testing.assert_array_equal(fib, U[k + "_fib"])
        assert st == US[k + "_stats"]
    fib, st = G.ibm_multi_direct_forcing_gpu(U["ibm_Lx"], U["ibm_Ly"], U["ibm_ud"], Uf,
                                             U["ibm_Ro"].ravel(), dx, dx, 1.0, La, ny, nx, 20, 5,
                                             dt, 6 * om, 1e-4)
    np.testing.assert_array_equal(fib, U[f"mdf_{dt}_div_fib"])
    assert st == US[f"mdf_{dt}_div_stats"]


@pytest.mark.parametrize("dt", ["hat", "peskin4pt"])
def test_dfc(dt):
    dx, La, ny, nx = US["ibm_dx"], US["ibm_Larea"], US["ibm_ny"], US["ibm_nx"]
    lam = G.compute_lambda_gpu(U["ibm_Lx"], U["ibm_Ly"], 1.0, La, dx, dx, ny, nx, dt)
    np.testing.assert_array_equal(lam, U[f"dfc_{dt}_lambda"])
    for given in (lam, None):
        fc, force, lam_used = G.apply_dfc_correction_gpu(U["ibm_Lx"], U["ibm_Ly"], U["ibm_ud"],
                                                         U[f"dfc_{dt}_fstar"], dx, dx, La, ny,
                                                         nx, dt, lambda_cache=given)
        np.testing.assert_array_equal(fc, U[f"dfc_{dt}_fcorr"])
        np.testing.assert_array_equal(force, U[f"dfc_{dt}_force"])
        np.testing.assert_array_equal(lam_used, U[f"dfc_{dt}_lambda"])
