"""Closed-form and brute-force checks of the oracle, independent of the golden vectors.

Each check is chosen so a plausible slip (dropped term, wrong sign or index, transposed
operand) fails it: conservation laws, exact special cases and brute-force sums.
"""
import numpy as np
import pytest

from oracle import oracle as O

E = np.array([[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1], [1, 1], [-1, 1], [-1, -1], [1, -1]],
             dtype=np.float64)
RNG = np.random.default_rng(11)


def test_feq_moments():
    ro = 1.0 + 0.1 * RNG.standard_normal(200)
    U = 0.05 * RNG.standard_normal((200, 2))
    f = O.feq(ro, U)
    np.testing.assert_allclose(f.sum(1), ro, rtol=1e-14)
    np.testing.assert_allclose(f @ E, ro[:, None] * U, atol=1e-15)
    # second moment: sum f e e = rho (cs2 I + u u)
    Pxx = (f * E[:, 0] ** 2).sum(1)
    Pxy = (f * E[:, 0] * E[:, 1]).sum(1)
    np.testing.assert_allclose(Pxx, ro * (1 / 3 + U[:, 0] ** 2), rtol=1e-13)
    np.testing.assert_allclose(Pxy, ro * U[:, 0] * U[:, 1], atol=1e-15)
    fi = O.feq(ro, U, True)  # incompressible: sum = rho, momentum = u
    np.testing.assert_allclose(fi.sum(1), ro, rtol=1e-14)
    np.testing.assert_allclose(fi @ E, U, atol=1e-15)


def test_macro_exact_on_integers():
    f = RNG.integers(0, 50, (30, 9)).astype(np.float64)
    ro, U = O.macroscopic(f, True)
    np.testing.assert_array_equal(ro, f.sum(1))
    np.testing.assert_array_equal(U, f @ E)


def test_stream_is_a_shift_brute_force():
    nx, ny = 8, 6
    f = np.arange(nx * ny * 9, dtype=np.float64).reshape(-1, 9)
    old = -np.ones_like(f)
    out = O.stream(old, f, nx, ny)
    for j in range(ny):
        for i in range(nx):
            for q in range(9):
                si, sj = i - int(E[q, 0]), j - int(E[q, 1])
                want = f[sj * nx + si, q] if 0 <= si < nx and 0 <= sj < ny else -1.0
                assert out[j * nx + i, q] == want


def test_bgk_conservation_and_tau1():
    N = 100
    ro = 1.0 + 0.05 * RNG.standard_normal(N)
    U = 0.05 * RNG.standard_normal((N, 2))
    feq = O.feq(ro, U)
    fs = feq + 1e-3 * RNG.standard_normal((N, 9))
    fib = 1e-3 * RNG.standard_normal((N, 2))
    zero = np.zeros((N, 2))
    # conserved mass and momentum when the moments of fs are those of feq
    fs_c = fs - (fs.sum(1) - ro)[:, None] / 9.0
    r2, u2 = O.macroscopic(fs_c)
    feq_c = O.feq(r2, u2)
    out = O.collide_bgk(fs_c, feq_c, u2, zero, 0.8)
    np.testing.assert_allclose(out.sum(1), r2, rtol=1e-14)
    np.testing.assert_allclose(out @ E, r2[:, None] * u2, atol=1e-14)
    # Guo forcing: momentum gain = (1 - 1/(2 tau)) * fib * dt, mass unchanged
    outF = O.collide_bgk(fs_c, feq_c, u2, fib, 0.8)
    np.testing.assert_allclose((outF - out).sum(1), 0.0, atol=1e-16)
    np.testing.assert_allclose((outF - out) @ E, (1 - 0.5 / 0.8) * fib, rtol=1e-9, atol=1e-16)
    # tau = 1: f = feq exactly (no force)
    np.testing.assert_allclose(O.collide_bgk(fs, feq, U, zero, 1.0), feq, atol=1e-16)


def test_trt_reduces_to_bgk():
    N = 64
    ro = 1.0 + 0.05 * RNG.standard_normal(N)
    U = 0.05 * RNG.standard_normal((N, 2))
    feq = O.feq(ro, U)
    fs = feq + 1e-3 * RNG.standard_normal((N, 9))
    fib = 1e-3 * RNG.standard_normal((N, 2))
    tau = 0.7
    a = O.collide_trt(fs, feq, U, fib, 1 / tau, 1 / tau)
    b = O.collide_bgk(fs, feq, U, fib, tau)
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-15)


def test_cm_mrt_identity_and_conservation():
    N = 64
    ro = 1.0 + 0.05 * RNG.standard_normal(N)
    U = 0.05 * RNG.standard_normal((N, 2))
    fs = O.feq(ro, U) + 1e-3 * RNG.standard_normal((N, 9))
    r, u = O.macroscopic(fs)
    zero = np.zeros((N, 2))
    # all rates zero: the central-moment round trip returns the input
    np.testing.assert_allclose(O.collide_cm(fs, u, zero, np.zeros(9)), fs, atol=2e-16)
    S = np.array([0, 0, 0, 1 / 0.7, 1 / 0.7, 1 / 0.7, 1.2, 1.2, 1.4])
    out = O.collide_cm(fs, u, zero, S)
    np.testing.assert_allclose(out.sum(1), r, rtol=1e-14)
    np.testing.assert_allclose(out @ E, r[:, None] * u, atol=1e-14)
    # all rates one: the output is the Maxwellian central-moment state; with u = 0 and a
    # unit density this is w (the D2Q9 weights)
    z = np.zeros((4, 2))
    one = O.collide_cm(np.tile(O.feq(np.ones(1), np.zeros((1, 2))), (4, 1)) +
                       1e-3 * np.array([[0, 1, -1, 1, -1, .5, -.5, .5, -.5]] * 4), z, z,
                       np.ones(9))
    w = O.feq(np.ones(1), np.zeros((1, 2)))[0]
    np.testing.assert_allclose(one, np.tile(w, (4, 1)), atol=1e-16)


@pytest.mark.parametrize("dt", ["hat", "peskin4pt"])
def test_delta_moments(dt):
    a = O.DELTA_RADIUS[dt]
    for x in RNG.uniform(-0.5, 0.5, 25):
        i = np.arange(-a - 1, a + 2)
        phi = O.delta(dt, i - x)
        assert abs(phi.sum() - 1.0) < 1e-14
        assert abs(((i - x) * phi).sum()) < 1e-14
    if dt == "peskin4pt":
        assert O.delta(dt, np.array([0.0, 1.0, 2.0, 2.5])).tolist() == [0.5, 0.25, 0.0, 0.0]


def _ring(n, r, cx, cy, dx):
    th = 2 * np.pi * np.arange(n) / n
    return (cx + r * np.cos(th)) * dx, (cy + r * np.sin(th)) * dx


@pytest.mark.parametrize("dt", ["hat", "peskin4pt"])
def test_df_spread_conserves_and_interp_is_exact_on_linear_fields(dt):
    ny, nx, dx = 30, 34, 1.0 / 29
    Lx, Ly = _ring(19, 6.3, 16.2, 14.7, dx)
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    Eux = 0.01 + 1e-3 * ii - 2e-3 * jj
    Euy = -0.02 + 5e-4 * jj
    Ro = np.ones((ny, nx))
    ud = 1e-3 * RNG.standard_normal((19, 2))
    Larea = 2 * np.pi * 6.3 / 19
    fib, Lux, Luy, Lfx, Lfy, R = O.direct_forcing(Lx, Ly, ud, Eux, Euy, Ro, dx, dx, 1.0, Larea,
                                                  ny, nx, dt)
    np.testing.assert_allclose(Lux, 0.01 + 1e-3 * Lx / dx - 2e-3 * Ly / dx, atol=1e-15)
    np.testing.assert_allclose(Luy, -0.02 + 5e-4 * Ly / dx, atol=1e-15)
    np.testing.assert_allclose(R, 1.0, atol=1e-15)
    np.testing.assert_allclose(fib.sum(0), [Lfx.sum() * Larea, Lfy.sum() * Larea], rtol=1e-12)


@pytest.mark.parametrize("dt", ["hat", "peskin4pt"])
def test_mdf_single_iteration_is_omega_times_df(dt):
    ny, nx, dx = 30, 34, 1.0 / 29
    Lx, Ly = _ring(19, 6.3, 16.2, 14.7, dx)
    U = 0.03 + 1e-3 * RNG.standard_normal((ny * nx, 2))
    ro = 1.0 + 1e-3 * RNG.standard_normal(ny * nx)
    ud = np.zeros((19, 2))
    Larea = 2 * np.pi * 6.3 / 19
    fib1 = O.direct_forcing(Lx, Ly, ud, U[:, 0].reshape(ny, nx), U[:, 1].reshape(ny, nx),
                            ro.reshape(ny, nx), dx, dx, 1.0, Larea, ny, nx, dt)[0]
    fibm, st = O.multi_direct_forcing(Lx, Ly, ud, U, ro, dx, dx, 1.0, Larea, ny, nx, 1, 1, dt,
                                      0.37, 1e-4)
    np.testing.assert_array_equal(fibm, fib1 * 0.37)
    assert st["iter_count"] == 1
    # the residual decreases monotonically for omega = 1/||A||_inf (Zhang 2020 Eq. 47)
    om = O.omega_from_type(Lx, Ly, dt, Larea, dx)
    res = [O.multi_direct_forcing(Lx, Ly, ud, U, ro, dx, dx, 1.0, Larea, ny, nx, n, n, dt, om,
                                  1e-30)[1]["last_residual"] for n in (1, 2, 3, 4, 5)]
    assert all(b < a for a, b in zip(res, res[1:]))


@pytest.mark.parametrize("dt", ["hat", "peskin4pt"])
def test_a_norm_separable_equals_dense_brute_force(dt):
    dx = 1.0 / 40
    Lx, Ly = _ring(11, 3.1, 20.3, 19.6, dx)
    Larea = 2 * np.pi * 3.1 / 11
    om = O.omega_from_type(Lx, Ly, dt, Larea, dx)
    # dense: A_lm = Larea * sum_x phi phi phi phi over a lattice patch
    X, Y = Lx / dx, Ly / dx
    gi = np.arange(10, 32)
    GX, GY = np.meshgrid(gi, gi, indexing="xy")
    phis = [O.delta(dt, GX - X[k]) * O.delta(dt, GY - Y[k]) for k in range(11)]
    A = np.array([[Larea * np.sum(phis[l] * phis[m]) for m in range(11)] for l in range(11)])
    assert abs(1.0 / A.sum(1).max() - om) < 1e-13 * om


@pytest.mark.parametrize("dt", ["hat", "peskin4pt"])
def test_dfc_lambda_isolated_marker(dt):
    ny, nx, dx = 20, 20, 1.0 / 19
    x, y = 9.37 * dx, 10.21 * dx
    lam = O.dfc_lambda(np.array([x]), np.array([y]), 1.0, 1.3, dx, dx, ny, nx, dt)
    a = O.DELTA_RADIUS[dt]
    i0, j0 = np.floor(x / dx + 0.5), np.floor(y / dx + 0.5)
    wx = O.delta(dt, x / dx - (i0 + np.arange(-a, a + 1)))
    wy = O.delta(dt, y / dx - (j0 + np.arange(-a, a + 1)))
    want = 1.0 / (2.0 * 1.3 * np.sum(wx ** 2) * np.sum(wy ** 2))
    assert abs(lam[0] - want) < 1e-13 * want


def test_dfc_bounce_back_of_equilibrium_reproduces_wall_velocity():
    """f_bb = f_opp + 2 w rho e.u/cs2: for an equilibrium at u_w the correction vanishes to
    O(u^3) (the bounce-back of f_eq(u_w) is f_eq(u_w) up to cubic terms)."""
    ny, nx, dx = 24, 24, 1.0 / 23
    uw = np.array([[0.002, -0.001]])
    U = np.tile(uw, (ny * nx, 1))
    fs = O.feq(np.ones(ny * nx), U)
    x, y = np.array([11.3 * dx]), np.array([12.6 * dx])
    lam = O.dfc_lambda(x, y, 1.0, 1.0, dx, dx, ny, nx, "hat")
    fc, force = O.dfc_correction(x, y, uw, fs, lam, dx, dx, 1.0, ny, nx, "hat")
    assert np.max(np.abs(fc - fs)) < 1e-8
    assert np.max(np.abs(force)) < 1e-8
