"""FP64 NumPy restatement of the learned MDF predictor's network, for tests only.

This is test infrastructure: it checks the device's forward, backward and Adam kernels and the
BF16 inference.  The network has no counterpart in the reference.  Layout of the parameter
vector (see include/iblbm_hip.h): W1[8][64] b1[64] W2[64][64] b2[64] W3[64][2] b3[2]; a layer
computes y_n = b_n + sum_k x_k W[k][n]; hidden activations are tanh; the loss is the mean of
(y_hat - y)^2 over rows and outputs.
"""
import numpy as np

IN, H, OUT = 8, 64, 2
SHAPES = [("W1", (IN, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, OUT)),
          ("b3", (OUT,))]
NP = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SHAPES])


def forward(p, x):
    w = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    h1 = np.tanh(x @ w["W1"] + w["b1"])
    h2 = np.tanh(h1 @ w["W2"] + w["b2"])
    y = h2 @ w["W3"] + w["b3"]
    return y, (x, h1, h2)


def loss_grad(p, x, y):
    w = unpack(p)
    yh, (x, h1, h2) = forward(p, x)
    r = yh - np.asarray(y, dtype=np.float64)
    loss = float(np.mean(r * r))
    d3 = 2.0 * r / r.size
    g = {"W3": h2.T @ d3, "b3": d3.sum(0)}
    d2 = (d3 @ w["W3"].T) * (1.0 - h2 * h2)
    g["W2"], g["b2"] = h1.T @ d2, d2.sum(0)
    d1 = (d2 @ w["W2"].T) * (1.0 - h1 * h1)
    g["W1"], g["b1"] = x.T @ d1, d1.sum(0)
    return loss, pack(g)


def bf16(a):
    """Round float32 values to BF16 (nearest even), returned as float64."""
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def forward_bf16(p, x):
    """The inference arithmetic: BF16 weights and activations, FP32-or-better accumulation."""
    w = unpack(p)
    h1 = np.tanh(bf16(x) @ bf16(w["W1"]) + w["b1"].astype(np.float32))
    h2 = np.tanh(bf16(h1) @ bf16(w["W2"]) + w["b2"].astype(np.float32))
    return bf16(h2) @ bf16(w["W3"]) + w["b3"].astype(np.float32)


# --------------------------------------------------------------------------- training

def norm_stats(rows):
    """The normalisation the device derives from its samples: feature mean and standard
    deviation (1 where the variance is below 1e-24), target root-mean-square (1 where below
    1e-30).  rows: (n, IN + OUT)."""
    r = np.asarray(rows, dtype=np.float64)
    n = max(r.shape[0], 1)
    mean = r.sum(0) / n
    var = (r * r).sum(0) / n - mean * mean
    sd = np.where(var[:IN] > 1e-24, np.sqrt(np.maximum(var[:IN], 0.0)), 1.0)
    rms = np.sqrt((r[:, IN:] ** 2).sum(0) / n)
    rms = np.where(rms > 1e-30, rms, 1.0)
    return np.concatenate([mean[:IN], sd, rms])


_M64 = (1 << 64) - 1


def pick(seed, s, b, n):
    """Row of minibatch slot b at training step s: splitmix64 of (seed << 32) ^ (1024 s + b),
    modulo n (the published SplitMix64 finaliser, Steele, Lea and Flood 2014)."""
    z = ((int(seed) << 32) ^ (int(s) * 1024 + int(b))) & _M64
    z = (z + 0x9E3779B97F4A7C15) & _M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    z ^= z >> 31
    return z % int(n)


def adam_step(p, m, v, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam update (Kingma and Ba 2015, Algorithm 1) at step t >= 1."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** t)
    vh = v / (1.0 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def train(p0, rows, norm, steps, lr, seed, batch=128, t0=0):
    """FP64 restatement of the device training loop: minibatches of `batch` rows picked by
    ``pick``, normalised by ``norm``, the loss gradient of ``loss_grad`` and Adam from zero
    moments.  Returns (parameters, minibatch losses)."""
    rows = np.asarray(rows, dtype=np.float64)
    n = rows.shape[0]
    p = np.asarray(p0, dtype=np.float64).copy()
    m, v = np.zeros_like(p), np.zeros_like(p)
    losses = []
    for s in range(steps):
        idx = [pick(seed, t0 + s, b, n) for b in range(batch)]
        mb = rows[idx]
        x = (mb[:, :IN] - norm[:IN]) / norm[IN:2 * IN]
        y = mb[:, IN:] / norm[2 * IN:]
        loss, g = loss_grad(p, x, y)
        losses.append(loss)
        p, m, v = adam_step(p, m, v, g, t0 + s + 1, lr)
    return p, np.array(losses)
