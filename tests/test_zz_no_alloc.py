"""The no-allocation MDF entry point (iblbm_op_multi_direct_forcing_noalloc) against the
allocating one (iblbm_op_multi_direct_forcing) through the C ABI: bit-identical fib and
statistics on two calls in a row, the second with a smaller grid and fewer markers, and the
limits refused with a clear error."""
import ctypes

import numpy as np
import pytest

from iblbm_amd import SimConfig, _abi
from iblbm_amd import gpu_kernels as G
from iblbm_amd.init import initialize

pytestmark = pytest.mark.gpu


def _case(NN, D_ratio, delta, seed, xmax=2.0):
    cfg = SimConfig(Re=40.0, NN=NN, xmax=xmax, ymax=1.0, cylinder_center=(0.55, 0.5),
                    cylinder_D_ratio=D_ratio, lattice_u=0.05, inflow_u=0.05,
                    bc_type="dirichlet_neumann", delta_type=delta, ibm_method="MDF",
                    collision_model="BGK", marker_spacing_factor=1.0)
    st = initialize(cfg, None)
    rng = np.random.default_rng(seed)
    N = st.nx * st.ny
    U = np.empty((N, 2))
    U[:, 0] = 0.05 * (1.0 + 0.2 * rng.uniform(-1, 1, N))
    U[:, 1] = 0.01 * rng.uniform(-1, 1, N)
    ro = 1.0 + 0.01 * rng.uniform(-1, 1, N)
    ud = 0.002 * rng.uniform(-1, 1, (st.Lb, 2))
    omega = G.estimate_optimal_omega_from_type(st.Lx, st.Ly, delta, st.Larea, st.dx)
    return st, U, ro, ud, omega


class _Hip:
    """Device buffers for the test through the HIP runtime the library itself links
    (libamdhip64.so.7): hipMalloc, hipMemcpy and hipFree by ctypes."""

    def __init__(self):
        _abi.load()  # loads the runtime as the library's dependency first
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                      ctypes.c_int]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.rt.hipDeviceSynchronize.argtypes = []
        self.live = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
        self.live.append(p)
        assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, shape):
        out = np.empty(shape)
        assert self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        assert self.rt.hipDeviceSynchronize() == 0
        for p in self.live:
            self.rt.hipFree(p)
        self.live = []


def _noalloc(st, U, ro, ud, delta, omega, n_iter=20, min_iter=5, tol=1e-4):
    hip = _Hip()
    d = {k: hip.put(v) for k, v in dict(Lx=st.Lx, Ly=st.Ly, ud=ud, U=U, ro=ro).items()}
    N = st.nx * st.ny
    fib = hip.put(np.full((N, 2), 7.0))  # stale values the call must overwrite
    stats = _abi.MdfStats()
    lib = _abi.load()
    rc = lib.iblbm_op_multi_direct_forcing_noalloc(
        d["Lx"], d["Ly"], d["ud"], st.Lb, d["U"], d["ro"], st.dx, st.dy, st.dt, st.Larea,
        st.ny, st.nx, n_iter, min_iter, _abi.DELTA[delta], omega, tol, fib,
        ctypes.byref(stats), None)
    msg = lib.iblbm_last_error().decode() if rc else ""
    out = hip.get(fib, (N, 2))
    hip.free()
    return rc, msg, out, stats


@pytest.mark.parametrize("delta", ["peskin4pt", "hat"])
def test_noalloc_matches_allocating_entry_point(delta):
    # first call: 241 x 121 grid, D = 0.4 (151 markers); second: 81 x 41, D = 0.25 (31)
    for NN, D_ratio, seed in ((121, 0.4, 1), (41, 0.25, 2)):
        st, U, ro, ud, omega = _case(NN, D_ratio, delta, seed)
        print(f"noalloc {delta} {st.nx}x{st.ny} Lb={st.Lb}", flush=True)
        ref, rst = G.ibm_multi_direct_forcing_gpu(st.Lx, st.Ly, ud, U, ro, st.dx, st.dy, st.dt,
                                                  st.Larea, st.ny, st.nx, n_iter=20, min_iter=5,
                                                  delta_type=delta, omega=omega, tol=1e-4)
        rc, msg, fib, sta = _noalloc(st, U, ro, ud, delta, omega)
        assert rc == 0, msg
        assert np.array_equal(fib, ref), float(np.max(np.abs(fib - ref)))
        assert sta.iter_count == rst["iter_count"]
        assert sta.last_residual == rst["last_residual"]
        assert bool(sta.diverged) == rst["diverged"] and bool(sta.converged) == rst["converged"]
        assert np.count_nonzero(fib) > 0


def test_noalloc_refuses_sizes_beyond_its_limits():
    # 513 markers: refused on the host before any launch
    st, U, ro, ud, omega = _case(161, 0.3, "hat", 3)
    Lx = np.concatenate([st.Lx] * 8)[:513]
    Ly = np.concatenate([st.Ly] * 8)[:513]
    big = type("S", (), dict(Lx=Lx, Ly=Ly, Lb=513, nx=st.nx, ny=st.ny, dx=st.dx, dy=st.dy,
                             dt=st.dt, Larea=st.Larea))
    rc, msg, _, _ = _noalloc(big, U, ro, np.zeros((513, 2)), "hat", omega)
    assert rc == _abi.EINVAL and "Lb = 513" in msg
    # a window box of 145 x 145 = 21025 > 16384 nodes (D = 0.5833 of a 241-node channel:
    # 440 markers): refused after the launch, output left zero
    st, U, ro, ud, omega = _case(241, 0.5833, "peskin4pt", 4, xmax=1.5)
    assert st.Lb == 440
    rc, msg, fib, _ = _noalloc(st, U, ro, ud, "peskin4pt", omega)
    assert rc == _abi.EINVAL and "exceeds 16384" in msg, msg
    assert not np.any(fib)
