"""Host-side pieces of the sedimentation driver (no GPU)."""
import numpy as np

from iblbm_amd.sedimentation import np_sum_sparse


def test_np_sum_sparse_matches_numpy_block_order():
    rng = np.random.default_rng(5)
    for n_total, start, n in ((300_001, 123_456, 9_001), (50_000, 8_190, 20_000),
                              (20_000, 0, 20_000), (8_192 * 3, 8_191, 2)):
        vals = rng.standard_normal(n) * np.exp(rng.standard_normal(n) * 4)
        full = np.zeros(n_total)
        full[start:start + n] = vals
        assert np_sum_sparse(vals, start, n_total) == float(np.sum(full))
        two = np.zeros((n_total, 2))
        two[start:start + n, 1] = vals
        assert np_sum_sparse(vals, start, n_total) == float(two[:, 1].sum())


def test_np_sum_sparse_matches_2d_sum():
    rng = np.random.default_rng(6)
    ny, nx = 193, 257
    y0, bh = 91, 37
    seg = rng.standard_normal(bh * nx)
    grid = np.zeros((ny, nx))
    grid[y0:y0 + bh, :] = seg.reshape(bh, nx)
    assert np_sum_sparse(seg, y0 * nx, ny * nx) == float(np.sum(grid))


def test_euler_free_fall_positions_follow_the_scheme():
    """No hydrodynamic force: the explicit Euler velocity grows by dt a each step with
    a = -g (rho_r - 1) / rho_r, and the position is x0 + dx dt^2 a n (n + 1) / 2 with the new
    velocity, x0 + dx dt^2 a n^2 / 2 with the trapezoidal mean (sums of the velocities)."""
    from iblbm_amd.sedimentation import euler_explicit
    rho_r, r, g, dt, dx = 1.5, 16.0, 2e-5, 1.0, 1.0 / 256
    mass = rho_r * np.pi * r**2
    a = -g * (rho_r - 1.0) / rho_r
    for scheme, tri in (("new_velocity", lambda n: n * (n + 1) / 2), ("trapezoidal",
                                                                     lambda n: n * n / 2)):
        pos, vel, prev = np.array([0.5, 3.0]), np.zeros(2), np.zeros(2)
        for n in range(1, 401):
            pos, v_new = euler_explicit("none", pos, vel, prev, np.zeros(2), mass, rho_r, r, g,
                                        dt, dx, scheme, "down")
            prev, vel = vel, v_new
        assert vel[0] == 0.0 and pos[0] == 0.5
        np.testing.assert_allclose(vel[1], 400 * dt * a, rtol=1e-12)
        np.testing.assert_allclose(pos[1] - 3.0, dx * dt * dt * a * tri(400), rtol=1e-10)


def test_euler_explicit_history_term():
    """With the hydrodynamic force balancing the net gravity the explicit-history model adds
    (m_f / m)(U^n - U^{n-1}) = (U^n - U^{n-1}) / rho_r; the none model adds nothing; gravity
    to the right acts on x."""
    from iblbm_amd.sedimentation import compute_net_gravity, euler_explicit
    rho_r, r, g = 1.25, 10.0, 3e-5
    mass = rho_r * np.pi * r**2
    for gdir in ("down", "right"):
        F = -compute_net_gravity(rho_r, r, g, gdir)
        v, vp = np.array([0.01, -0.02]), np.array([0.004, -0.005])
        _, vn = euler_explicit("explicit_history", np.zeros(2), v, vp, F, mass, rho_r, r, g,
                               1.0, 1.0, "new_velocity", gdir)
        np.testing.assert_allclose(vn, v + (v - vp) / rho_r, rtol=1e-13)
        _, vn0 = euler_explicit("none", np.zeros(2), v, vp, F, mass, rho_r, r, g, 1.0, 1.0,
                                "new_velocity", gdir)
        np.testing.assert_allclose(vn0, v, rtol=1e-13)
    _, vr = euler_explicit("none", np.zeros(2), np.zeros(2), np.zeros(2), np.zeros(2), mass,
                           rho_r, r, g, 1.0, 1.0, "new_velocity", "right")
    assert vr[0] > 0.0 and vr[1] == 0.0


def test_euler_rotation_under_constant_torque():
    """Constant torque T from rest: omega_n = n dt T / I, the angle dt^2 T / I n (n + 1) / 2
    (new rate) or n^2 / 2 (trapezoidal); an increment of the internal angular momentum adds
    dL / I to the rate."""
    from iblbm_amd.sedimentation import euler_rotation
    T, I, dt = 2.5e-3, 7.0e3, 1.0
    for scheme, tri in (("new_velocity", lambda n: n * (n + 1) / 2), ("trapezoidal",
                                                                     lambda n: n * n / 2)):
        om, ang = 0.0, 0.0
        for _ in range(300):
            om, inc = euler_rotation(om, T, I, dt, 0.0, scheme)
            ang += inc
        np.testing.assert_allclose(om, 300 * dt * T / I, rtol=1e-12)
        np.testing.assert_allclose(ang, dt * dt * T / I * tri(300), rtol=1e-10)
    om, _ = euler_rotation(0.0, 0.0, I, dt, 0.35, "new_velocity")
    np.testing.assert_allclose(om, 0.35 / I, rtol=1e-15)


def test_internal_angular_momentum_closed_forms():
    """compute_l_int_z's sum on a box: a solid-body rotation u = Omega x r at unit density
    gives Omega times the sum of r^2 over the nodes inside the disc (summed here node by node
    in a plain loop), a uniform flow about a centre on a node gives zero (the disc's nodes
    are symmetric), and a disc at the grid's corner reads the box clipped to the grid."""
    from iblbm_amd.sedimentation import l_int_box, l_int_z
    nx, ny, dx = 120, 90, 0.01
    for cx, cy, r in ((0.503, 0.417, 0.123), (0.6, 0.45, 0.2), (0.04, 0.03, 0.1)):
        i_lo, i_hi, j_lo, j_hi = l_int_box(nx, ny, dx, cx, cy, r)
        assert 0 <= i_lo < i_hi <= nx and 0 <= j_lo < j_hi <= ny
        ci, cj, rl = cx / dx, cy / dx, r / dx
        ii, jj = np.meshgrid(np.arange(i_lo, i_hi, dtype=float),
                             np.arange(j_lo, j_hi, dtype=float))
        Om = 3e-4
        U = np.stack([-Om * (jj - cj), Om * (ii - ci)], axis=-1)
        ro = np.ones_like(ii)
        got = l_int_z(ro, U, i_lo, j_lo, dx, cx, cy, r)
        ref = 0.0
        n_in = 0
        for j in range(0, ny):
            for i in range(0, nx):
                rx, ry = i - ci, j - cj
                if rx * rx + ry * ry <= rl * rl:
                    ref += Om * (rx * rx + ry * ry)
                    n_in += 1
        assert n_in > 0
        np.testing.assert_allclose(got, ref, rtol=1e-12)
    # uniform flow, centre on a node: the sum over the symmetric disc vanishes
    cx, cy, r = 0.5, 0.4, 0.15
    i_lo, i_hi, j_lo, j_hi = l_int_box(nx, ny, dx, cx, cy, r)
    h, w = j_hi - j_lo, i_hi - i_lo
    U = np.zeros((h, w, 2))
    U[..., 0], U[..., 1] = 0.02, -0.01
    scale = 0.03 * (r / dx) * np.pi * (r / dx) ** 2  # |U| |r| times the disc's nodes
    assert abs(l_int_z(np.ones((h, w)), U, i_lo, j_lo, dx, cx, cy, r)) < 1e-14 * scale
    # corner: the box is clipped at 0
    assert l_int_box(nx, ny, dx, 0.04, 0.03, 0.1)[0::2] == (0, 0)
