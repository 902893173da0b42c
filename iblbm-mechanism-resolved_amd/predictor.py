"""The learned MDF force predictor of a device-resident MDF instance.

Modes: "off" (plain MDF), "collect" (plain MDF that records, per marker and step, the
features and the target, both scaled: the change since the previous step of plain MDF's force
before its last sub-step, so that a learned step's one true correction lands on plain MDF's
own result, truncation included), "warm" (the previous step's force before its last sub-step
is the starting guess, no network), "learned" (that force plus the network's predicted
change).  The features are the change of the forcing slip q = R (ud - U) before
forcing since the previous step, the direction of q, the marker direction from the body
centre and the sub-cell offset; changes are divided by sig, the largest change of q over the
64 markers of the marker's wavefront, so the prediction keeps its accuracy in scale while the
changes decay along a run.  A sample is recorded only when the previous step ran plain MDF.
In the last two modes every step still applies at least one true MDF correction, keeps
iterating until plain MDF's tolerance is met, and restarts plain MDF once if the divergence
guard trips.  Training runs on the device (FP32 forward, backward and Adam); inference is BF16
on the MFMA units inside the window workgroup.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _abi
from ._abi import check, f32, fptr


class Predictor:
    def __init__(self, ctx, k: int):
        self.ctx, self.k = ctx, k

    def mode(self, mode: str, capacity: int = 1 << 20):
        if mode not in _abi.PRED:
            raise KeyError(f"predictor mode {mode!r} (expected one of {sorted(_abi.PRED)})")
        check(self.ctx.lib.iblbm_predictor_mode(self.ctx.h, self.k, _abi.PRED[mode],
                                                int(capacity)), "predictor_mode")

    def samples(self) -> int:
        n = ctypes.c_int64()
        check(self.ctx.lib.iblbm_predictor_samples(self.ctx.h, self.k, ctypes.byref(n)),
              "predictor_samples")
        return int(n.value)

    def samples_written(self) -> int:
        n = ctypes.c_int64()
        check(self.ctx.lib.iblbm_predictor_samples_written(self.ctx.h, self.k, ctypes.byref(n)),
              "predictor_samples_written")
        return int(n.value)

    def data(self, first: int, n: int) -> np.ndarray:
        """Samples first .. first + n - 1 as rows (features[8], target[2])."""
        rows = np.empty((int(n), _abi.MLP_IN + _abi.MLP_OUT), dtype=np.float32)
        check(self.ctx.lib.iblbm_predictor_data(self.ctx.h, self.k, int(first), int(n),
                                                fptr(rows)), "predictor_data")
        return rows

    def train(self, steps: int = 2000, lr: float = 1e-3, seed: int = 1) -> np.ndarray:
        losses = np.empty(int(steps), dtype=np.float32)
        check(self.ctx.lib.iblbm_predictor_train(self.ctx.h, self.k, int(steps), float(lr),
                                                 int(seed) & 0xFFFFFFFF, fptr(losses)),
              "predictor_train")
        return losses

    def params(self):
        p = np.empty(_abi.MLP_NP, dtype=np.float32)
        nrm = np.empty(_abi.MLP_NORM, dtype=np.float32)
        check(self.ctx.lib.iblbm_predictor_params(self.ctx.h, self.k, fptr(p), fptr(nrm)),
              "predictor_params")
        return p, nrm

    def set_params(self, params, norm):
        p, nrm = f32(params), f32(norm)
        if p.size != _abi.MLP_NP or nrm.size != _abi.MLP_NORM:
            raise ValueError("predictor parameters have the wrong size")
        check(self.ctx.lib.iblbm_predictor_set_params(self.ctx.h, self.k, fptr(p), fptr(nrm)),
              "predictor_set_params")

    def previous_force(self, Lb: int) -> np.ndarray:
        """The previous step's marker forces before its last MDF sub-step, then its forcing
        slips q, (2 Lb, 2)."""
        F = np.empty((2 * int(Lb), 2))
        check(self.ctx.lib.iblbm_predictor_previous_force(self.ctx.h, self.k, _abi.dptr(F), 0),
              "predictor_previous_force")
        return F

    def set_previous_force(self, F):
        F = _abi.f64(F).reshape(-1, 2)
        check(self.ctx.lib.iblbm_predictor_previous_force(self.ctx.h, self.k, _abi.dptr(F), 1),
              "predictor_previous_force")

    def stats(self, reset: bool = False) -> dict:
        out = np.zeros(4)
        check(self.ctx.lib.iblbm_predictor_stats(self.ctx.h, self.k, _abi.dptr(out), int(reset)),
              "predictor_stats")
        steps = max(out[1], 1.0)
        return {"substeps_per_step": out[0] / steps, "steps": int(out[1]),
                "first_residual": out[2] / steps, "restarts": int(out[3])}


def mlp_loss_grad(params, x, y):
    """Device FP32 loss and gradient of one batch (<= 128 rows, normalised x and y)."""
    p, x, y = f32(params), f32(x), f32(y)
    g = np.empty(_abi.MLP_NP, dtype=np.float32)
    loss = np.empty(1, dtype=np.float32)
    check(_abi.load().iblbm_mlp_loss_grad(fptr(p), fptr(x), fptr(y), int(x.shape[0]), fptr(g),
                                          fptr(loss)), "mlp_loss_grad")
    return float(loss[0]), g


def mlp_infer_bf16(params, x):
    """Device BF16 MFMA inference of n rows of normalised features."""
    p, x = f32(params), f32(x)
    out = np.empty((x.shape[0], _abi.MLP_OUT), dtype=np.float32)
    check(_abi.load().iblbm_mlp_infer_bf16(fptr(p), fptr(x), int(x.shape[0]), fptr(out)),
          "mlp_infer_bf16")
    return out
