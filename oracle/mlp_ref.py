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
