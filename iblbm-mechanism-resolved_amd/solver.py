"""Time loop of the reference (``iblbm/runtime/loop.py:44-151``) on the fused gfx950 step.

``run(cfg)`` keeps the reference's call surface and result dictionary.  The device owns the
populations; one launch per step runs streaming, the boundary closure, the moments, the
immersed-boundary stage and the collision; the host only reads FP64 reductions at the
diagnostic interval.

Step order (``iblbm/runtime/loop.py:59-87``): the reference collides once before the loop and
then runs stream -> boundary -> macroscopic -> IBM -> (velocity correction) -> feq -> collide
for DF/MDF, and collide -> stream -> boundary -> macroscopic -> DFC -> macroscopic -> feq for
DFC.  Both are the same sequence once the first collision is hoisted out of the loop, which
is what the device runs: the buffer between launches always holds post-collision f.
"""
from __future__ import annotations

import ctypes
import math
import time

import numpy as np

from . import _abi, gpu_kernels
from ._abi import check, dptr, f64
from .init import initialize


class Context:
    """One HIP stream and the instances stepped together on it (``iblbm_ctx_*``)."""

    def __init__(self, device: int = 0):
        self.lib = _abi.load()
        h = ctypes.c_void_p()
        check(self.lib.iblbm_ctx_create(int(device), ctypes.byref(h)), "ctx_create")
        self.h = h
        self.n = []  # node counts per instance

    def close(self):
        if self.h:
            self.lib.iblbm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, case: _abi.Case) -> int:
        i = ctypes.c_int()
        check(self.lib.iblbm_instance_add(self.h, ctypes.byref(case), ctypes.byref(i)),
              "instance_add")
        self.n.append((case.nx, case.ny))
        return i.value

    def set_markers(self, k, Lx, Ly, ud, Larea):
        Lx, Ly, ud = f64(Lx), f64(Ly), f64(ud).reshape(-1, 2)
        check(self.lib.iblbm_set_markers(self.h, k, dptr(Lx), dptr(Ly), dptr(ud), Lx.shape[0],
                                         float(Larea)), "set_markers")

    def set_boundary_velocity(self, k, U):
        """The previous velocity the first step's boundary closure reads (N, 2)."""
        U = f64(U).reshape(-1, 2)
        check(self.lib.iblbm_set_boundary_velocity(self.h, k, dptr(U)), "set_boundary_velocity")

    def set_analytic_boundary(self, k, coef, decay):
        """Analytic boundary tables: per boundary node (A, B), and d(t) for t = 0, 1, ..."""
        coef, decay = f64(coef).reshape(-1, 2), f64(decay).reshape(-1)
        check(self.lib.iblbm_set_analytic_boundary(self.h, k, dptr(coef), coef.shape[0],
                                                   dptr(decay), decay.shape[0]),
              "set_analytic_boundary")

    def set_desired_velocity(self, k, ud):
        ud = f64(ud).reshape(-1, 2)
        check(self.lib.iblbm_set_desired_velocity(self.h, k, dptr(ud)), "set_desired_velocity")

    def set_torque_center(self, k, cx, cy):
        check(self.lib.iblbm_set_torque_center(self.h, k, float(cx), float(cy)), "torque_center")

    def upload_f(self, k, f):
        f = f64(f).reshape(-1, 9)
        check(self.lib.iblbm_upload_f(self.h, k, dptr(f)), "upload_f")

    def download_f(self, k):
        nx, ny = self.n[k]
        out = np.empty((nx * ny, 9))
        check(self.lib.iblbm_download_f(self.h, k, dptr(out)), "download_f")
        return out

    def download_state(self, k):
        nx, ny = self.n[k]
        N = nx * ny
        fs, ro, U = np.empty((N, 9)), np.empty(N), np.empty((N, 2))
        check(self.lib.iblbm_download_state(self.h, k, dptr(fs), dptr(ro), dptr(U)),
              "download_state")
        return fs, ro, U

    def download_state_box(self, k, x0, y0, w, h):
        """rho (h, w) and U (h, w, 2) of the last step on the box [x0, x0 + w) x
        [y0, y0 + h): the same values as ``download_state``'s ro and U there."""
        ro, U = np.empty((h, w)), np.empty((h, w, 2))
        check(self.lib.iblbm_download_state_box(self.h, k, int(x0), int(y0), int(w), int(h),
                                                dptr(ro), dptr(U)), "download_state_box")
        return ro, U

    def download_fib(self, k):
        nx, ny = self.n[k]
        out = np.empty((nx * ny, 2))
        check(self.lib.iblbm_download_fib(self.h, k, dptr(out)), "download_fib")
        return out

    def download_dfc_force(self, k, Lb):
        out = np.empty((Lb, 2))
        check(self.lib.iblbm_download_dfc_force(self.h, k, dptr(out)), "download_dfc_force")
        return out

    def download_window(self, k):
        """(fib_x, fib_y) on the window box (box-major) and the box [x0, y0, bw, bh]."""
        nx, ny = self.n[k]
        fx, fy = np.empty(nx * ny), np.empty(nx * ny)
        box = (ctypes.c_int * 4)()
        check(self.lib.iblbm_download_window(self.h, k, dptr(fx), dptr(fy), box),
              "download_window")
        nb = box[2] * box[3]
        return fx[:nb], fy[:nb], (box[0], box[1], box[2], box[3])

    def step(self, n, diag_every=0, graph=False):
        fn = self.lib.iblbm_graph_run if graph else self.lib.iblbm_step
        check(fn(self.h, int(n), int(diag_every)), "step")

    def step_persistent(self, n, diag_every=0):
        """n steps in one persistent launch of the fused step kernel (grid barrier between
        steps); returns when they are done.  Needs the fused launch, one collision model, the
        tile row layout and no boundary kernel (ValueError otherwise)."""
        check(self.lib.iblbm_step_persistent(self.h, int(n), int(diag_every)), "step_persistent")

    def sync(self):
        check(self.lib.iblbm_sync(self.h), "sync")

    def step_counter(self) -> int:
        """Steps run by the context so far (diagnostics fire on multiples of diag_every)."""
        return int(self.lib.iblbm_step_counter(self.h))

    def step_mode(self) -> str:
        """"fused" (one launch per step), "two_streams" or "serial"."""
        m = ctypes.c_int()
        check(self.lib.iblbm_step_mode(self.h, ctypes.byref(m)), "step_mode")
        return ("fused", "two_streams", "serial")[m.value]

    def schedule(self) -> dict:
        """The launch schedule (``iblbm_get_schedule``): mode and bulk work distribution."""
        sc = _abi.Schedule()
        check(self.lib.iblbm_get_schedule(self.h, ctypes.byref(sc)), "get_schedule")
        mode = {v: k for k, v in _abi.MODES.items()}[sc.mode]
        return {"mode": mode, "bulk_threads": sc.bulk_threads, "bulk_rows": sc.bulk_rows,
                "bulk_waves": sc.bulk_waves, "step_rows": sc.step_rows,
                "step_order": {-1: "queue", 0: "stride", 1: "blocked"}.get(sc.step_order,
                                                                          sc.step_order),
                "row_layout": ROW_LAYOUTS[sc.row_layout],
                "stores": STORE_MODES[sc.store_mode], "bulk_delay_us": sc.bulk_delay / 100.0}

    def set_schedule(self, mode="auto", bulk_threads=256, bulk_rows=1, bulk_waves=0,
                     step_rows=2, step_order="stride", row_layout="tiles", stores="normal",
                     bulk_delay_us=0.0):
        """step_order: "stride", "blocked" (one block per wavefront), a block size k >= 2, or
        "queue" (one block per workgroup, its wavefronts taking rows from a counter in LDS);
        row_layout: "tiles" (64-column tiles of lattice rows) or "flat" (64 consecutive
        row-major nodes per bulk row); stores: "normal" or "nontemporal"; bulk_delay_us: the
        fused launch's bulk start delay (0 to 10 us, in steps of 10 ns)."""
        order = {"queue": -1, "stride": 0, "blocked": 1}.get(step_order, step_order)
        sc = _abi.Schedule(_abi.MODES[mode], bulk_threads, bulk_rows, bulk_waves, step_rows,
                           int(order), ROW_LAYOUTS.index(row_layout),
                           STORE_MODES.index(stores), int(round(bulk_delay_us * 100)))
        check(self.lib.iblbm_set_schedule(self.h, ctypes.byref(sc)), "set_schedule")

    def row_layout_in_use(self) -> str:
        """The bulk row layout the instances use now ("tiles" or "flat")."""
        v = ctypes.c_int()
        check(self.lib.iblbm_row_layout_in_use(self.h, ctypes.byref(v)), "row_layout_in_use")
        return ROW_LAYOUTS[v.value]

    def autotune(self, steps: int = 20) -> dict:
        """Times the candidate schedules on the current state (which is restored) and keeps
        the fastest; returns it with its time per step."""
        ms, n = ctypes.c_double(), ctypes.c_int()
        check(self.lib.iblbm_autotune(self.h, int(steps), ctypes.byref(ms), ctypes.byref(n)),
              "autotune")
        out = self.schedule()
        out.update(ms_per_step=ms.value, candidates=n.value,
                   row_layout_in_use=self.row_layout_in_use())
        return out

    def init_flow(self, k, rho, ux, uy, amp=0.0, seed=0):
        """Seeded synthetic equilibrium state on the device (``iblbm_init_flow``)."""
        check(self.lib.iblbm_init_flow(self.h, k, float(rho), float(ux), float(uy), float(amp),
                                       int(seed) & 0xFFFFFFFFFFFFFFFF), "init_flow")

    def set_concurrent(self, on: bool):
        check(self.lib.iblbm_set_concurrent(self.h, int(bool(on))), "set_concurrent")

    def mdf_stats(self, k):
        st = _abi.MdfStats()
        check(self.lib.iblbm_get_mdf_stats(self.h, k, ctypes.byref(st)), "mdf_stats")
        return {"iter_count": st.iter_count, "last_residual": st.last_residual,
                "diverged": bool(st.diverged), "converged": bool(st.converged),
                "guard_trips": st.guard_trips, "omega_used": st.omega_used}

    def set_particles(self, k, lb, larea):
        """Several particles in one DF / MDF window (markers concatenated in this order)."""
        lb = np.ascontiguousarray(lb, dtype=np.int32)
        la = f64(larea)
        check(self.lib.iblbm_set_particles(self.h, k, len(lb),
                                           lb.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                           dptr(la)), "set_particles")

    def download_particle_fib(self, k, part):
        """Coupled MDF: particle ``part``'s accumulated window fib and the window box."""
        nx, ny = self.n[k]
        fx, fy = np.empty(nx * ny), np.empty(nx * ny)
        box = (ctypes.c_int * 4)()
        check(self.lib.iblbm_download_particle_fib(self.h, k, int(part), dptr(fx), dptr(fy), box),
              "download_particle_fib")
        nb = box[2] * box[3]
        return fx[:nb], fy[:nb], (box[0], box[1], box[2], box[3])

    def download_marker_force(self, k, Lb):
        """DF of several particles: the marker forces (Lb, 2) of the last step."""
        F = np.empty((int(Lb), 2))
        check(self.lib.iblbm_download_marker_force(self.h, k, dptr(F)), "download_marker_force")
        return F

    def set_markers_batch(self, ids, lb, Lx, Ly, ud, larea, center=None):
        """``set_markers`` (and ``set_torque_center``) of several instances in one call:
        instance ``ids[b]`` takes the next ``lb[b]`` markers of the concatenated arrays."""
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        lb = np.ascontiguousarray(lb, dtype=np.int32)
        Lx, Ly, ud, la = f64(Lx), f64(Ly), f64(ud), f64(larea)
        n, tot = len(ids), int(lb.sum())
        if len(lb) != n or len(la) != n or Lx.size != tot or Ly.size != tot or ud.size != 2 * tot:
            raise ValueError("set_markers_batch: array sizes do not match the marker counts")
        cp = None
        if center is not None:
            c = f64(center)
            if c.size != 2 * n:
                raise ValueError("set_markers_batch: center needs (n, 2) entries")
            cp = dptr(c)
        ip = ctypes.POINTER(ctypes.c_int)
        check(self.lib.iblbm_set_markers_batch(self.h, n, ids.ctypes.data_as(ip),
                                               lb.ctypes.data_as(ip), dptr(Lx), dptr(Ly),
                                               dptr(ud), dptr(la), cp), "set_markers_batch")

    def download_marker_forces_batch(self, ids, lb_total):
        """DFC forces q (or several-particle DF marker forces) of several instances,
        concatenated in ``ids`` order: (lb_total, 2)."""
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        F = np.empty((int(lb_total), 2))
        check(self.lib.iblbm_download_marker_forces_batch(
            self.h, len(ids), ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), dptr(F),
            int(lb_total)), "download_marker_forces_batch")
        return F

    def diag(self, k, slot=None):
        """Diagnostics of the last diagnostic step; with ``slot``, those fetched into that
        pinned slot by ``diag_fetch`` (waits for that copy only)."""
        d = _abi.Diag()
        if slot is None:
            check(self.lib.iblbm_get_diag(self.h, k, ctypes.byref(d)), "get_diag")
        else:
            check(self.lib.iblbm_get_diag_slot(self.h, k, int(slot), ctypes.byref(d)),
                  "get_diag_slot")
        return d

    def diag_fetch(self, slot):
        """Queue the copy of the last diagnostic step's sums into pinned slot 0 or 1."""
        check(self.lib.iblbm_diag_fetch(self.h, int(slot)), "diag_fetch")

    def mark(self, k):
        """Record timing mark k (0..3) on the context stream."""
        check(self.lib.iblbm_event_record(self.h, int(k)), "event_record")

    def elapsed_ms(self, a, b):
        ms = ctypes.c_double()
        check(self.lib.iblbm_event_elapsed(self.h, int(a), int(b), ctypes.byref(ms)),
              "event_elapsed")
        return ms.value

    def last_timing(self):
        ms, n = ctypes.c_double(), ctypes.c_int()
        check(self.lib.iblbm_last_timing(self.h, ctypes.byref(ms), ctypes.byref(n)), "timing")
        return ms.value, n.value


def bc_name(cfg) -> str:
    """The boundary closure a configuration runs: ``tg_analytical`` for the Taylor-Green
    scenario (``iblbm/boundary/__init__.py:62-75``), else ``bc_type``."""
    return "tg_analytical" if cfg.scenario_type == "taylor_green" else cfg.bc_type


def ibm_name(cfg):
    """The immersed-boundary method a configuration runs (none for a Taylor-Green vortex
    without its cylinder, ``iblbm/ibm/dispatch.py:38-39``)."""
    if cfg.scenario_type == "taylor_green" and not cfg.tg_with_ibm:
        return None
    return cfg.ibm_method


def make_case(cfg, st, precision=_abi.FP64) -> _abi.Case:
    """Translate a ``SimConfig`` + initial state into the library's instance description."""
    if bc_name(cfg) not in _abi.BC:
        raise KeyError(f"bc_type={cfg.bc_type!r} is outside the GPU hot path "
                       f"(supported: {sorted(_abi.BC)})")
    if cfg.ibm_method not in ("DF", "MDF", "DFC"):
        raise KeyError(f"unknown ibm_method={cfg.ibm_method!r} (registered: ['DF', 'DFC', 'MDF'])")
    if cfg.delta_type not in _abi.DELTA:
        raise KeyError(cfg.delta_type)
    # stream_first (runtime/loop.py:60-87) moves the collision from the top of a step to its
    # end, after one collision before the loop: the sequence of operations is the same, and
    # no hook between the end of one step and the top of the next changes what the collision
    # reads, so it runs unchanged (the reference's stream-first runs equal its default-order
    # runs bit for bit)
    c = _abi.Case()
    c.nx, c.ny = st.nx, st.ny
    c.collision = _abi.COLL[cfg.collision_model]
    c.bc = _abi.BC[bc_name(cfg)]
    c.ibm = _abi.IBM[ibm_name(cfg)]
    c.delta = _abi.DELTA[cfg.delta_type]
    c.incompressible = int(bool(cfg.incompressible_lbgk))
    c.precision = precision
    c.tau = st.tau
    c.tau_minus = st.tau_minus if st.tau_minus is not None else 1.0
    S = st.S_diag if st.S_diag is not None else np.zeros(9)
    for q in range(9):
        c.s_diag[q] = float(S[q])
    c.dt, c.dx, c.dy = st.dt, st.dx, st.dy
    c.in_u, c.dens = st.in_u, st.dens
    c.mdf_iterations, c.mdf_min_iterations = cfg.mdf_iterations, cfg.mdf_min_iterations
    c.mdf_tolerance = cfg.mdf_tolerance
    c.mdf_omega = 1.0
    a = _abi.DELTA_RADIUS[cfg.delta_type]
    span = int(math.ceil(2.0 * st.lattice_r)) + 2 * a + 4
    c.lb_capacity = max(st.Lb, 1)
    c.box_capacity = span * span
    c.fixed_body = 1 if cfg.motion_type is None else 0
    return c


def pre_collide(cfg, st):
    """The collision before the loop (``iblbm/runtime/loop.py:59-61``): f from fstar = feq."""
    fib = np.zeros((st.nodenums, 2))
    if cfg.collision_model == "BGK":
        return gpu_kernels.collision_gpu(st.feq0, st.feq0, st.U0, fib, st.tau, st.dt)
    if cfg.collision_model == "TRT":
        return gpu_kernels.collision_trt_gpu(st.feq0, st.feq0, st.U0, fib, st.tau, st.tau_minus,
                                             st.dt)
    return gpu_kernels.collision_cm_mrt_gpu(st.feq0, st.feq0, st.U0, fib, st.S_diag, st.dt)


def oscillating_markers(t, Lx_c, Ly_c, lattice_u, KC, r, lattice_r):
    """``update_oscillating`` (``iblbm/physics/motion.py:15-52``): x(t) = A sin(2 pi f0 t)
    with A = KC D / (2 pi), f0 = U / (KC D); the markers move by -x(t) and their desired
    velocity is (x(t), 0)."""
    f0 = lattice_u / (KC * 2.0 * lattice_r)
    A = KC * 2.0 * r / (2.0 * np.pi)
    xt = A * np.sin(f0 * 2.0 * np.pi * t)
    Lx = Lx_c - xt
    Ly = Ly_c.copy()
    ud = np.empty((len(Lx_c), 2))
    ud[:, 0] = xt
    ud[:, 1] = 0.0
    return Lx, Ly, ud


PRECISIONS = {"fp64": _abi.FP64, "fp32dev": _abi.FP32_DEV}
ROW_LAYOUTS = ["tiles", "flat"]
STORE_MODES = ["normal", "nontemporal"]


class FixedCylinder:
    """A fixed cylinder on the device, stepped in chunks.

    precision "fp64" reproduces the reference's FP64 arithmetic bit for bit; "fp32dev" stores
    the populations as FP32 deviations from the weights and keeps the immersed-boundary stage
    (interpolation, forces, MDF residual, reductions) in FP64.
    """

    def __init__(self, cfg, perturb=None, device=0, ctx=None, precision="fp64", init="host"):
        if cfg.motion_type not in (None, "oscillating"):
            raise ValueError("FixedCylinder runs motion_type None or 'oscillating'; use the "
                             "sedimentation driver")
        if precision not in PRECISIONS:
            raise KeyError(f"precision={precision!r} (expected one of {sorted(PRECISIONS)})")
        if init not in ("host", "device"):
            raise KeyError(f"init={init!r} (expected 'host' or 'device')")
        self.cfg = cfg
        self.precision = precision
        # "host": the reference's initial state (numpy perturbation, pre-loop collision);
        # "device": the same equilibrium formed on the device from a hash-seeded perturbation,
        # for ensembles too large to set up on the host
        self.st = initialize(cfg, perturb if init == "host" else None, fields=init == "host")
        self.ctx = ctx or Context(device)
        case = make_case(cfg, self.st, PRECISIONS[precision])
        self.omega = None
        self.has_ibm = ibm_name(cfg) is not None
        if self.has_ibm and cfg.ibm_method == "MDF":
            self.omega = gpu_kernels.estimate_optimal_omega_from_type(
                self.st.Lx, self.st.Ly, cfg.delta_type, self.st.Larea, self.st.dx)
            case.mdf_omega = self.omega
        self.k = self.ctx.add(case)
        self.Lx_c, self.Ly_c = self.st.Lx.copy(), self.st.Ly.copy()
        if self.has_ibm:
            self.ctx.set_markers(self.k, self.st.Lx, self.st.Ly, self.st.desired_velocity,
                                 self.st.Larea)
            cx, cy = cfg.cylinder_center
            self.ctx.set_torque_center(self.k, cx, cy)
        if init == "host":
            self.ctx.upload_f(self.k, pre_collide(cfg, self.st))
        else:
            seed, amp = perturb if perturb is not None else (0, 0.0)
            self.ctx.init_flow(self.k, 1.0, cfg.inflow_u, 0.0, amp, seed)
        name = bc_name(cfg)
        if name in ("velocity_inlet", "open_boundary"):  # the first step's previous velocity
            self.ctx.set_boundary_velocity(self.k, self.st.U0)
        elif name == "tg_analytical":
            from . import taylor_green as TG
            self.ctx.set_analytic_boundary(self.k, TG.boundary_coefficients(cfg, self.st.nx,
                                                                            self.st.ny),
                                           TG.decay_table(cfg, max(int(cfg.max_steps), 1)))
        # host work between steps: the Taylor-Green markers' desired velocity
        # (taylor_green.py:48-49, before the step) and the oscillating markers
        # (fixed_cylinder.py:55-66, after it)
        self.tg_markers = name == "tg_analytical" and self.has_ibm
        self.oscillating = cfg.motion_type == "oscillating"
        self._markers_due = None
        self.ttt = 0

    def advance(self, n, diag_every=0, graph=False, persistent=False):
        """n steps: launched per step (replayed as a graph with graph=True) or, with
        persistent=True, as one persistent launch (fixed bodies on the fused path)."""
        if n <= 0:
            return
        if not (self.tg_markers or self.oscillating):
            if persistent:
                self.ctx.step_persistent(n, diag_every)
            else:
                self.ctx.step(n, diag_every, graph)
            self.ttt += n
            return
        if persistent:
            raise ValueError("moving markers are set between steps: no persistent launch")
        cfg, st = self.cfg, self.st
        for _ in range(n):  # one step at a time, with the host's marker updates around it
            if self.tg_markers:
                from . import taylor_green as TG
                ud = TG.desired_velocity(cfg, st.Lx, st.Ly, st.dx, st.dy, st.nx, st.ny,
                                         (self.ttt + 1) * st.dt)
                self.ctx.set_desired_velocity(self.k, ud)
            if self._markers_due is not None:  # the positions computed after the last step
                self.ctx.set_markers(self.k, st.Lx, st.Ly, self._markers_due, st.Larea)
                self._markers_due = None
            self.ctx.step(1, diag_every, False)
            self.ttt += 1
            if self.oscillating:
                # the host's positions move now, as the reference's do; the device's marker
                # window keeps the step's until the next step, so that the state and fib read
                # between steps are those of the step just taken
                st.Lx, st.Ly, self._markers_due = oscillating_markers(
                    self.ttt, self.Lx_c, self.Ly_c, cfg.lattice_u, cfg.KC, st.r, st.lattice_r)

    @property
    def predictor(self):
        """The learned MDF force predictor of this instance (``iblbm_amd.predictor``)."""
        from .predictor import Predictor
        return Predictor(self.ctx, self.k)

    def cd_cl(self, slot=None):
        """Cd, Cl of the last diagnostic step (``iblbm/diagnostics/core.py:53-113``), or of
        the one fetched into ``slot``."""
        d = self.ctx.diag(self.k, slot)
        ro_mean = d.rho_sum / self.st.nodenums
        # fixed_cylinder.py:69-71: an oscillating cylinder is scaled by lattice_u
        u_ref = self.cfg.lattice_u if self.cfg.motion_type == "oscillating" else self.st.in_u
        denom = ro_mean * u_ref ** 2 * self.st.lattice_r
        return float(abs(d.force[0]) / denom), float(d.force[1] / denom)


class LoopControl:
    """Termination bookkeeping of the reference's time loop (``iblbm/runtime/loop.py:64-70``,
    ``:113-137``), apart from the stepping so that it can be driven on the host alone.

    * The convergence exit is tested before the step bound at the top of each pass, so a run
      that converges at its last check ends "converged".
    * The velocity reference field is taken only at the check step equal to
      ``convergence_start``; a start that is not a check step never converges.
    * The Cd-window criterion sets "cd_converged" (``loop.py:132-140``).
    """

    def __init__(self, cfg):
        self.cfg = cfg
        self.converged = False
        self.termination = "max_steps"
        self.error = 1.0
        self.pre_U = None

    def done(self, ttt: int) -> bool:
        cfg = self.cfg
        if self.converged and (cfg.use_convergence or cfg.cd_convergence_tol is not None):
            if self.termination == "max_steps":
                self.termination = "converged"
            return True
        return ttt >= cfg.max_steps

    def check(self, ttt: int, Cd_hist, velocity):
        """At a check step: ``velocity()`` returns the (N, 2) field, read only when needed."""
        cfg = self.cfg
        if cfg.use_convergence and (ttt == cfg.convergence_start or
                                    (ttt > cfg.convergence_start and self.pre_U is not None)):
            U = np.asarray(velocity())
            if ttt == cfg.convergence_start:
                self.pre_U = U.copy()
            else:  # iblbm/diagnostics/core.py:116-137
                pre = self.pre_U
                numer = np.sum((U[:, 0] - pre[:, 0]) ** 2 + (U[:, 1] - pre[:, 1]) ** 2)
                denom = np.sum(pre[:, 0] ** 2 + pre[:, 1] ** 2)
                self.error = 1.0 if denom == 0.0 else float(np.sqrt(numer / denom))
                self.pre_U = U.copy()
                if self.error < cfg.convergence_threshold:
                    self.converged = True
        if (cfg.cd_convergence_tol is not None and ttt > cfg.cd_convergence_start_step
                and len(Cd_hist) >= cfg.cd_convergence_window):
            recent = Cd_hist[-cfg.cd_convergence_window:]
            if max(recent) - min(recent) < cfg.cd_convergence_tol:
                self.converged = True
                self.termination = "cd_converged"


def run(cfg, verbose: bool = True, callback=None, save_dir=None, perturb=None,
        device: int = 0, graph: bool = True, precision: str = "fp64") -> dict:
    """``iblbm.run`` for the fixed-cylinder scenario on the fused GPU step.

    Returns the reference's result keys: Cd_history, Cl_history, converged, final_step,
    termination_reason, plus ``mdf_stats`` (last step of each diagnostic interval) and
    ``sim`` (the device-resident simulation, whose state can be downloaded).
    """
    if cfg.motion_type == "sedimentation":
        raise ValueError(f"motion_type={cfg.motion_type!r}: sedimentation runs through "
                         "iblbm_amd.sedimentation.run")
    sim = FixedCylinder(cfg, perturb=perturb, device=device, precision=precision)
    mon = None
    if save_dir:  # status.json in the reference's format
        from .monitor import FixedCylinderMonitor
        mon = FixedCylinderMonitor(save_dir, cfg)
    ci = int(cfg.check_interval)
    Cd_hist, Cl_hist, mdf_hist = [], [], []
    ctl = LoopControl(cfg)
    t0 = time.time()
    while not ctl.done(sim.ttt):
        n = min(ci - sim.ttt % ci, cfg.max_steps - sim.ttt)
        sim.advance(n, diag_every=ci, graph=graph)
        ttt = sim.ttt
        if ttt % ci != 0:
            continue
        Cd, Cl = sim.cd_cl()
        Cd_hist.append(Cd)
        Cl_hist.append(Cl)
        if cfg.ibm_method == "MDF":
            mdf_hist.append(sim.ctx.mdf_stats(sim.k))
        if verbose:
            print(f"step {ttt:>7d} | Cd={Cd:.6f} Cl={Cl:.6f} | err={ctl.error:.2e} | "
                  f"{time.time() - t0:.1f}s")
        if callback is not None:
            callback(step=ttt, Cd=Cd, Cl=Cl, error=ctl.error, state=sim,
                     elapsed=time.time() - t0, converged=ctl.converged)
        if mon is not None:
            mon(ttt, Cd, Cl, error=ctl.error, converged=ctl.converged, elapsed=time.time() - t0)
        ctl.check(ttt, Cd_hist, lambda: sim.ctx.download_state(sim.k)[2])
    converged, termination = ctl.converged, ctl.termination
    sim.ctx.sync()
    if mon is not None:
        mon.finalize(sim.ttt, termination, converged=converged, elapsed=time.time() - t0)
    return {"Cd_history": np.array(Cd_hist), "Cl_history": np.array(Cl_hist),
            "converged": converged, "final_step": sim.ttt, "termination_reason": termination,
            "mdf_stats": mdf_hist, "sim": sim, "state": sim.st}
