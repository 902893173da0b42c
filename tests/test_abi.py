"""C-ABI library: loads, exports every declared symbol, host-side helpers (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, golden
import iblbm_amd
from iblbm_amd import _abi, gpu_kernels

HEADER = os.path.join(ROOT, "include", "iblbm_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^[A-Za-z_][\w \t\*]*?\b(iblbm_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = _abi.load()
    syms = declared_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    assert set(syms) == set(_abi.SIGNATURES), set(syms) ^ set(_abi.SIGNATURES)


def test_case_struct_layout_matches_header():
    # 8 int32 + 2 doubles + 9 doubles + 5 doubles + 2 int32 + 2 doubles + 4 int32
    assert ctypes.sizeof(_abi.Case) == 8 * 4 + 16 * 8 + 2 * 4 + 2 * 8 + 4 * 4
    assert ctypes.sizeof(_abi.MdfStats) == 4 * 4 + 2 * 8
    assert ctypes.sizeof(_abi.Diag) == 48


@pytest.mark.parametrize("dt", ["hat", "peskin4pt"])
def test_omega_estimate_matches_reference(dt):
    U, US = golden("units")
    om = gpu_kernels.estimate_optimal_omega_from_type(U["ibm_Lx"], U["ibm_Ly"], dt,
                                                      US["ibm_Larea"], US["ibm_dx"])
    assert om == US[f"omega_{dt}"]


def test_omega_estimate_of_golden_runs():
    for case in ("run_mdf_trt_p4", "run_mdf_bgk_hat"):
        g, sc = golden(case)
        from conftest import golden_cfg
        from iblbm_amd.init import initialize
        cfg = golden_cfg(sc)
        st = initialize(cfg)
        np.testing.assert_array_equal(st.Lx, g["Lx"])
        np.testing.assert_array_equal(st.Ly, g["Ly"])
        om = gpu_kernels.estimate_optimal_omega_from_type(st.Lx, st.Ly, cfg.delta_type,
                                                          st.Larea, st.dx)
        assert om == sc["omega"]


def test_marker_box_covers_clamped_stencils():
    lib = _abi.load()
    rng = np.random.default_rng(3)
    nx, ny, dx = 40, 30, 1.0 / 29
    for delta, a in ((0, 2), (1, 1)):
        Lx = rng.uniform(-0.1, (nx + 2) * dx, 17)
        Ly = rng.uniform(-0.1, (ny + 2) * dx, 17)
        box = (ctypes.c_int * 4)()
        _abi.check(lib.iblbm_marker_box(_abi.dptr(Lx), _abi.dptr(Ly), 17, dx, dx, nx, ny, delta,
                                        box))
        ix0 = np.floor(Lx / dx + 0.5).astype(np.int64)
        iy0 = np.floor(Ly / dx + 0.5).astype(np.int64)
        xs = np.clip(np.concatenate([ix0 + d for d in range(-a, a + 1)]), 0, nx - 1)
        ys = np.clip(np.concatenate([iy0 + d for d in range(-a, a + 1)]), 0, ny - 1)
        assert (box[0], box[1]) == (xs.min(), ys.min())
        assert (box[0] + box[2] - 1, box[1] + box[3] - 1) == (xs.max(), ys.max())


def test_errors_map_to_reference_exception_types():
    lib = _abi.load()
    box = (ctypes.c_int * 4)()
    x = np.zeros(1)
    with pytest.raises(ValueError):
        _abi.check(lib.iblbm_marker_box(_abi.dptr(x), _abi.dptr(x), 0, 0.1, 0.1, 4, 4, 0, box))
    with pytest.raises(KeyError):
        gpu_kernels.estimate_optimal_omega_from_type(x, x, "gaussian", 1.0, 0.1)


def test_config_validation_mirrors_reference():
    from iblbm_amd import SimConfig
    with pytest.raises(ValueError):
        SimConfig(Re=20, NN=21, mdf_iterations=0)
    with pytest.raises(ValueError):
        SimConfig(Re=20, NN=21, motion_type="sedimentation")  # inertia model unset
    msg = SimConfig(Re=100, NN=641, cylinder_D_ratio=1 / 16, lattice_u=0.04,
                    collision_model="TRT").warn_if_unstable()
    assert msg is not None and "0.5480" in msg


def test_host_initial_state_matches_reference_bits():
    from conftest import RUN_CASES, golden_cfg, perturb_of
    from iblbm_amd.init import initialize
    from oracle import oracle as O
    for case in RUN_CASES:
        g, sc = golden(case)
        cfg = golden_cfg(sc)
        st = initialize(cfg, perturb_of(sc))
        ref = O.initialize(cfg, perturb_of(sc))
        np.testing.assert_array_equal(st.feq0, ref.feq)
        np.testing.assert_array_equal(st.U0, ref.U)
        assert st.tau == sc["tau"] and st.Larea == sc["Larea"] and st.Lb == sc["Lb"]
        if sc.get("tau_minus") is not None:
            assert st.tau_minus == sc["tau_minus"]


def test_gpu_requested_without_device_fails_loudly():
    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK):
        pytest.skip("a GPU device node is present")
    from iblbm_amd.solver import Context
    with pytest.raises(RuntimeError):
        Context(0)


# Every name the reference imports from its gpu_kernels module (iblbm/ibm/dispatch.py:21,
# mdf.py:131, multi_particle.py:112, forcing.py:57-63 and :157-164, df.py:81, dfc.py:296 and
# :307, lbm/collision/{bgk,trt,cm_mrt}.py, lbm/streaming.py:83, lbm/equilibrium.py:48,
# lbm/macroscopic.py:56), with the parameter names and defaults of its definition
# (iblbm/gpu_kernels.py:751-1147).
REFERENCE_OPERATORS = {
    "_grid": (["n"], {}),
    "collision_gpu": (["fstar", "feq", "U", "fib", "tau", "dt"], {}),
    "collision_trt_gpu": (["fstar", "feq", "U", "fib", "tau", "tau_minus", "dt"], {}),
    "collision_cm_mrt_gpu": (["fstar", "feq", "U", "fib", "S_diag", "dt"], {}),
    "streaming_gpu": (["fstar_old", "f", "nx", "ny"], {}),
    "macroscopic_gpu": (["fstar", "incompressible"], {"incompressible": False}),
    "feq_gpu": (["ro", "U", "incompressible"], {"incompressible": False}),
    "velocity_correction_gpu": (["U", "fib", "ro", "dt", "incompressible"],
                                {"incompressible": False}),
    "ibm_direct_forcing_gpu": (["Lx", "Ly", "desired_vel", "Eux", "Euy", "Ro", "dx", "dy", "dt",
                                "Larea", "ny", "nx", "delta_type", "return_interp_vel",
                                "incompressible"],
                               {"delta_type": "peskin4pt", "return_interp_vel": False,
                                "incompressible": False}),
    "ibm_direct_forcing_fields_gpu": (["Lx", "Ly", "desired_vel", "Eux", "Euy", "Ro", "dx", "dy",
                                       "dt", "Larea", "ny", "nx", "delta_type",
                                       "incompressible"],
                                      {"delta_type": "peskin4pt", "incompressible": False}),
    "compute_lambda_gpu": (["Lx", "Ly", "rho_f", "Larea", "dx", "dy", "ny", "nx", "delta_type"],
                           {}),
    "dfc_correction_gpu": (["Lx", "Ly", "desired_vel", "fstar", "lambda_k", "dx", "dy", "Larea",
                            "ny", "nx", "delta_type", "lattice"], {}),
}
REFERENCE_CONSTANTS = {"_BLOCK": 256, "_DELTA_TYPE_IDS": {"peskin4pt": 0, "hat": 1},
                       "_DELTA_STENCIL_A": {"peskin4pt": 2, "hat": 1}}


def test_every_reference_import_exists_with_its_signature():
    import inspect
    from iblbm_amd import gpu_kernels as G
    for name, (params, defaults) in REFERENCE_OPERATORS.items():
        sig = inspect.signature(getattr(G, name))
        assert list(sig.parameters) == params, (name, list(sig.parameters))
        got = {k: p.default for k, p in sig.parameters.items()
               if p.default is not inspect.Parameter.empty}
        assert got == defaults, (name, got)
    for name, value in REFERENCE_CONSTANTS.items():
        assert getattr(G, name) == value, name
    assert G._grid(257) == (2,) and G._grid(256) == (1,)
    for name in ("_ibm_interp_force_kern", "_ibm_spread_kern"):
        assert callable(getattr(G, name))
