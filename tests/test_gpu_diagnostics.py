"""Boundary slip error: the device interpolation against the C oracle (bit-exact) and a
closed form (a uniform field interpolates to itself)."""
import numpy as np
import pytest

from iblbm_amd.diagnostics import compute_slip_error, cylinder_markers, interpolate_to_markers
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("delta_type", ["hat", "peskin4pt"])
def test_interpolation_matches_oracle(delta_type):
    rng = np.random.default_rng(5)
    NN, ny, nx = 33, 33, 81
    dx = dy = 1.0 / (NN - 1)
    Eux, Euy = rng.standard_normal((ny, nx)), rng.standard_normal((ny, nx))
    Lx, Ly, _ = cylinder_markers(1.0, 0.5, 0.2, NN)
    Lux, Luy = interpolate_to_markers(Eux, Euy, Lx, Ly, dx, dy, delta_type)
    Lb = len(Lx)
    _, ox, oy, *_ = O.direct_forcing(Lx, Ly, np.zeros((Lb, 2)), Eux, Euy, np.ones((ny, nx)),
                                     dx, dy, 1.0, 1.0, ny, nx, delta_type)
    np.testing.assert_array_equal(Lux, ox)
    np.testing.assert_array_equal(Luy, oy)


@pytest.mark.parametrize("delta_type", ["hat", "peskin4pt"])
def test_uniform_field_slip(delta_type):
    NN, ny, nx = 33, 33, 81
    dx = dy = 1.0 / (NN - 1)
    a, b, uref = 0.03, -0.04, 0.05
    res = compute_slip_error(np.full((ny, nx), a), np.full((ny, nx), b), dx, dy, 1.0, 0.5, 0.4,
                             NN, delta_type, u_ref=uref)
    # sum of the kernel weights is 1 (to round-off), so every marker sees (a, b)
    assert res["mean"] == pytest.approx(np.hypot(a, b) / uref, rel=1e-13)
    assert res["max"] == pytest.approx(np.hypot(a, b) / uref, rel=1e-13)
    t = res["theta"]
    ux_t, uy_t = np.full(len(t), a), np.full(len(t), b)
    res0 = compute_slip_error(np.full((ny, nx), a), np.full((ny, nx), b), dx, dy, 1.0, 0.5, 0.4,
                              NN, delta_type, u_target_x=ux_t, u_target_y=uy_t, u_ref=uref)
    assert res0["max"] < 1e-13
