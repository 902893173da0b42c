"""The settling driver's force and torque sums: NumPy's summation of a full-grid vector that
is zero outside a window, computed from the window's values only (``WindowSums``), equals
``np.sum`` of the materialised vector bit for bit, for windows anywhere on the grid,
including those whose last reduction buffer is cut short by the end of the vector."""
import numpy as np

from iblbm_amd.sedimentation import WindowSums


def test_window_sums_equal_full_vector_sums():
    rng = np.random.default_rng(11)
    for nx, ny in ((1281, 3841), (257, 1921), (97, 300), (64, 64)):
        N = nx * ny
        ws = WindowSums(nx, N)
        box = None
        for trial in range(40):
            if trial % 4 == 0:
                bw = int(rng.integers(4, min(180, nx) + 1))
                bh = int(rng.integers(4, min(180, ny) + 1))
                x0 = int(rng.integers(0, nx - bw + 1))
                y0 = ny - bh if trial % 12 == 0 else int(rng.integers(0, ny - bh + 1))
                box = (x0, y0, bw, bh)
            x0, y0, bw, bh = box
            vals = [rng.standard_normal(bw * bh) * (rng.random(bw * bh) < 0.4)
                    for _ in range(3)]
            got = ws.sums(box, *vals)
            for v, g in zip(vals, got):
                full = np.zeros(N)
                full.reshape(ny, nx)[y0:y0 + bh, x0:x0 + bw] = v.reshape(bh, bw)
                assert float(np.sum(full)) == g
            assert ws.sums(box, vals[0], vals[1]) == got[:2]


def test_planted_order_change_is_caught():
    """Summing the window's values directly (another order) differs from NumPy's full-grid
    order for some window: the test above would catch a shortcut of that kind."""
    rng = np.random.default_rng(2)
    nx, ny = 1281, 600
    diffs = 0
    for _ in range(20):
        v = rng.standard_normal(150 * 150)
        full = np.zeros(nx * ny)
        full.reshape(ny, nx)[200:350, 500:650] = v.reshape(150, 150)
        diffs += float(np.sum(full)) != float(np.sum(v))
    assert diffs > 0
