"""The predictor network's CPU oracle, pinned by finite differences and BF16 rounding facts."""
import numpy as np

from oracle import mlp_ref as M


def _rand(seed=0, rows=16):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(M.NP) * 0.3
    return p, rng.standard_normal((rows, M.IN)), rng.standard_normal((rows, M.OUT))


def test_gradient_is_the_derivative_of_the_loss():
    p, x, y = _rand()
    loss, g = M.loss_grad(p, x, y)
    rng = np.random.default_rng(1)
    idx = rng.choice(M.NP, 60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros(M.NP)
        e[i] = h
        fd = (M.loss_grad(p + e, x, y)[0] - M.loss_grad(p - e, x, y)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-6 * abs(g[i]), (i, fd, g[i])
    assert loss == np.mean((M.forward(p, x)[0] - y) ** 2)


def test_pack_unpack_layout():
    p = np.arange(M.NP, dtype=float)
    w = M.unpack(p)
    assert w["W1"][0, 1] == 1 and w["b1"][0] == M.IN * M.H
    assert w["W2"][1, 0] == M.IN * M.H + M.H + M.H and w["b3"][-1] == M.NP - 1
    np.testing.assert_array_equal(M.pack(w), p)


def test_bf16_rounding():
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)  # BF16 keeps 8 significant bits
    vals = np.array([one, one + ulp / 2, one + 3 * ulp / 2, -2.5, 3.0e38], dtype=np.float32)
    # ties go to even: 1 + ulp/2 -> 1, 1 + 3 ulp/2 -> 1 + 2 ulp
    np.testing.assert_array_equal(M.bf16(vals)[:4], [1.0, 1.0, 1.0 + 2 * ulp, -2.5])
    assert np.isfinite(M.bf16(vals)[4])
