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


def test_pick_is_splitmix64():
    # the first SplitMix64 output from state 0 is 0xE220A8397B1DCDAF (published sequence)
    assert M.pick(0, 0, 0, 1 << 64) == 0xE220A8397B1DCDAF
    # slot b and step s enter as 1024 s + b
    assert M.pick(5, 3, 7, 1 << 64) == M.pick(5, 0, 3 * 1024 + 7, 1 << 64)
    rows = [M.pick(11, s, b, 1000) for s in range(20) for b in range(128)]
    assert min(rows) >= 0 and max(rows) < 1000 and len(set(rows)) > 900


def test_adam_constant_gradient_steps_by_lr():
    # with a constant gradient the bias-corrected moments equal g and g^2 at every step, so
    # each update is -lr g / (|g| + eps)
    p, m, v = np.zeros(3), np.zeros(3), np.zeros(3)
    g = np.array([2.0, -0.5, 1e-3])
    for t in range(1, 6):
        q, m, v = M.adam_step(p, m, v, g, t, lr=0.01)
        np.testing.assert_allclose(q - p, -0.01 * g / (np.abs(g) + 1e-8), rtol=1e-9)
        p = q


def test_adam_minimises_a_quadratic():
    p, m, v = np.array([3.0, -2.0]), np.zeros(2), np.zeros(2)
    for t in range(1, 3001):
        p, m, v = M.adam_step(p, m, v, np.array([1.0, 10.0]) * p, t, lr=0.01)
    assert np.all(np.abs(p) < 1e-2)


def test_norm_stats_closed_form():
    rows = np.zeros((4, M.IN + M.OUT))
    rows[:, 0] = [1.0, 2.0, 3.0, 4.0]     # mean 2.5, sd sqrt(1.25)
    rows[:, 1] = 7.0                      # constant: sd -> 1
    rows[:, M.IN] = [3.0, -3.0, 3.0, -3.0]  # target rms 3
    nrm = M.norm_stats(rows)
    assert nrm[0] == 2.5 and abs(nrm[M.IN] - np.sqrt(1.25)) < 1e-15
    assert nrm[1] == 7.0 and nrm[M.IN + 1] == 1.0
    assert nrm[2 * M.IN] == 3.0 and nrm[2 * M.IN + 1] == 1.0  # all-zero target -> 1


def test_train_restatement_reduces_the_loss():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((512, M.IN))
    rows = np.concatenate([x, np.stack([np.sin(x[:, 0]), x[:, 1] * x[:, 2]], 1)], 1)
    p0 = rng.standard_normal(M.NP) * 0.2
    p, losses = M.train(p0, rows, M.norm_stats(rows), steps=200, lr=3e-3, seed=4)
    assert losses[-20:].mean() < 0.5 * losses[:20].mean()
