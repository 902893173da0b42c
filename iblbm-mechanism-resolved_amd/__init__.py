"""iblbm_amd -- the gfx950 (MI355X) IB-LBM per-step hot path.

Public surface (mirrors the reference package ``iblbm``):
  - ``SimConfig``      run description (``iblbm/config.py``)
  - ``run``            time loop on the fused GPU step (``iblbm/runtime/loop.py``)
  - ``gpu_kernels``    operator-level functions with the reference's wrapper signatures

Backend selection follows ``iblbm/backend.py:14-36``: ``IBLBM_GPU`` (default ``"1"``) requests
the GPU, and a missing library or device raises ``RuntimeError`` instead of falling back.
"""
from __future__ import annotations

import os

from .config import SimConfig

_gpu_requested = os.environ.get("IBLBM_GPU", "1") != "0"


def run(cfg, verbose: bool = True, callback=None, save_dir=None, **kw):
    """``iblbm.run`` on the GPU hot path (fixed cylinder or sedimentation)."""
    if not _gpu_requested:
        raise RuntimeError("IBLBM_GPU=0 selects the reference's CPU backend, which this "
                           "package does not provide; unset it to use the GPU path")
    if cfg.motion_type == "sedimentation":
        from .sedimentation import run as _run
    else:
        from .solver import run as _run
    return _run(cfg, verbose=verbose, callback=callback, save_dir=save_dir, **kw)


__all__ = ["SimConfig", "run"]
