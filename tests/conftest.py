import json
import os
import subprocess
import sys
from types import SimpleNamespace

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: larger parity cases")
    # the C oracle and the HIP library are built in-tree; build only what is missing
    if not os.path.exists(os.path.join(ROOT, "oracle", "libcpu_ref.so")):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                       capture_output=True)
    lib = os.path.join(ROOT, "iblbm-mechanism-resolved_amd", "libiblbm_hip.so")
    if not os.path.exists(lib):
        import iblbm_amd.build as b  # noqa: E402
        b.build()


def golden(name):
    arr = np.load(os.path.join(GOLDEN, name + ".npz"))
    with open(os.path.join(GOLDEN, name + ".json")) as fh:
        sc = json.load(fh)
    return arr, sc


def golden_cfg(sc):
    from iblbm_amd import SimConfig
    return SimConfig(**sc["cfg"])


def perturb_of(sc):
    return (sc["perturb_seed"], sc["perturb_amp"]) if "perturb_seed" in sc else None


RUN_CASES = ["run_df_bgk_hat", "run_mdf_trt_p4", "run_dfc_cm_hat", "run_df_cm_p4",
             "run_dfc_trt_p4", "run_mdf_bgk_hat", "run_sf_mdf_trt_p4", "run_sf_dfc_cm_hat"]
# runs with the velocity-inlet, open and Taylor-Green boundaries (oscillating cylinder and
# Taylor-Green vortex scenarios)
BC_RUN_CASES = ["run_vi_df_bgk_hat", "run_vi_mdf_cm_p4", "run_osc_df_bgk_hat",
                "run_osc_mdf_trt_p4", "run_osc_dfc_cm_hat", "run_tg_lbm_d10", "run_tg_df_d10",
                "run_tg_mdf_d10", "run_tg_dfc_d10"]


def rel_linf(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
