"""The operator-level module bound the way the reference's own GPU branches call it, against
the reference's vectors: each test replays one call site with the reference's argument order
and arithmetic around the call, bit for bit.

* ``iblbm/ibm/dfc.py:296-316``: lambda from ``compute_lambda_gpu`` (or the cache), then
  ``fstar + dfc_correction_gpu(...)[0]``.
* ``iblbm/ibm/dispatch.py:18-33``: ``velocity_correction_gpu`` in place on the post-stream
  velocity of a whole run's last step.
* ``iblbm/ibm/forcing.py:57-99, :157-247``: the private kernel symbols with ``_grid`` /
  ``_BLOCK`` launch tuples, outputs in caller arrays, ``Efx``/``Efy`` zeroed by the caller.
* ``iblbm/ibm/mdf.py:131-140`` and ``multi_particle.py:112-128``: ``ibm_direct_forcing_fields_gpu``
  with and without ``incompressible``.
"""
import numpy as np
import pytest

from conftest import golden
from iblbm_amd import gpu_kernels as G

pytestmark = pytest.mark.gpu
U, US = golden("units")


@pytest.mark.parametrize("dt", ["hat", "peskin4pt"])
def test_dfc_call_site(dt):
    dx, La, ny, nx = US["ibm_dx"], US["ibm_Larea"], US["ibm_ny"], US["ibm_nx"]
    Lx, Ly, ud, fstar = U["ibm_Lx"], U["ibm_Ly"], U["ibm_ud"], U[f"dfc_{dt}_fstar"]
    for cached in (None, U[f"dfc_{dt}_lambda"]):
        if cached is not None:
            lambda_k = cached
        else:
            lambda_k = G.compute_lambda_gpu(Lx, Ly, rho_f=1.0, Larea=La, dx=dx, dy=dx, ny=ny,
                                            nx=nx, delta_type=dt)
        delta_f_euler, dfc_force = G.dfc_correction_gpu(Lx, Ly, ud, fstar, lambda_k, dx, dx, La,
                                                        ny, nx, delta_type=dt, lattice=None)
        assert delta_f_euler.shape == fstar.shape and dfc_force.shape == (len(Lx), 2)
        fstar_corrected = fstar + delta_f_euler
        np.testing.assert_array_equal(fstar_corrected, U[f"dfc_{dt}_fcorr"])
        np.testing.assert_array_equal(dfc_force, U[f"dfc_{dt}_force"])
        # the field is the correction alone: zero away from the markers
        assert np.count_nonzero(np.any(delta_f_euler != 0.0, axis=1)) < 0.5 * len(fstar)


@pytest.mark.parametrize("case", ["run_df_bgk_hat", "run_mdf_trt_p4", "run_df_cm_p4",
                                  "run_mdf_bgk_hat"])
def test_velocity_correction_call_site(case):
    g, sc = golden(case)
    ro, Upre = G.macroscopic_gpu(g["fstar"], False)
    np.testing.assert_array_equal(ro, g["ro"])
    state_U = Upre.copy()
    out = G.velocity_correction_gpu(state_U, g["fib"], ro, 1.0, incompressible=False)
    assert out is None  # in place, as the reference's wrapper
    np.testing.assert_array_equal(state_U, g["U"])
    # incompressible form: u + 0.5 fib dt
    inc = Upre.copy()
    G.velocity_correction_gpu(inc, g["fib"], ro, 1.0, incompressible=True)
    np.testing.assert_array_equal(inc, Upre + 0.5 * g["fib"] * 1.0)


@pytest.mark.parametrize("dt", ["hat", "peskin4pt"])
@pytest.mark.parametrize("inc", [False, True])
def test_private_marker_kernels_call_site(dt, inc):
    dx, La, ny, nx = US["ibm_dx"], US["ibm_Larea"], US["ibm_ny"], US["ibm_nx"]
    Lx, Ly, ud = U["ibm_Lx"], U["ibm_Ly"], U["ibm_ud"]
    Eux, Euy, Ro = U["ibm_Eux"], U["ibm_Euy"], U["ibm_Ro"]
    Lb = len(Lx)
    tid, a = G._DELTA_TYPE_IDS[dt], G._DELTA_STENCIL_A[dt]
    Lfx, Lfy, Lux, Luy = (np.empty(Lb) for _ in range(4))
    G._ibm_interp_force_kern(G._grid(Lb), (G._BLOCK,),
                             (Lx, Ly, ud, Eux, Euy, Ro, np.float64(dx), np.float64(dx),
                              np.float64(1.0), np.int32(nx), np.int32(ny), np.int32(Lb),
                              np.int32(tid), np.int32(a), np.int32(int(inc)), Lfx, Lfy, Lux, Luy))
    k = f"df_{dt}_{'i' if inc else 'c'}"
    np.testing.assert_array_equal(Lfx, U[k + "_Lfx"])
    np.testing.assert_array_equal(Lfy, U[k + "_Lfy"])
    np.testing.assert_array_equal(Lux, U[k + "_Lux"])
    np.testing.assert_array_equal(Luy, U[k + "_Luy"])
    Efx, Efy = np.zeros((ny, nx)), np.zeros((ny, nx))
    G._ibm_spread_kern(G._grid(Lb), (G._BLOCK,),
                       (Lfx, Lfy, Lx, Ly, np.float64(dx), np.float64(dx), np.float64(La),
                        np.int32(nx), np.int32(ny), np.int32(Lb), np.int32(tid), np.int32(a),
                        Efx, Efy))
    fib_p = np.empty((nx * ny, 2))
    fib_p[:, 0] = Efx.ravel()
    fib_p[:, 1] = Efy.ravel()
    np.testing.assert_array_equal(fib_p, U[k + "_fib"])


@pytest.mark.parametrize("dt", ["hat", "peskin4pt"])
def test_direct_forcing_fields_call_site(dt):
    dx, La, ny, nx = US["ibm_dx"], US["ibm_Larea"], US["ibm_ny"], US["ibm_nx"]
    for inc in (False, True):
        Efx, Efy, Lux, Luy = G.ibm_direct_forcing_fields_gpu(
            U["ibm_Lx"], U["ibm_Ly"], U["ibm_ud"], U["ibm_Eux"], U["ibm_Euy"], U["ibm_Ro"], dx,
            dx, 1.0, La, ny, nx, delta_type=dt, incompressible=inc)
        k = f"df_{dt}_{'i' if inc else 'c'}"
        np.testing.assert_array_equal(Efx.ravel(), U[k + "_fib"][:, 0])
        np.testing.assert_array_equal(Efy.ravel(), U[k + "_fib"][:, 1])
        np.testing.assert_array_equal(Lux, U[k + "_Lux"])
        np.testing.assert_array_equal(Luy, U[k + "_Luy"])
