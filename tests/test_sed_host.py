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
