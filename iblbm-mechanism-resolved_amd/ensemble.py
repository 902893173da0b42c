"""Case matrices and seeded ensembles stepped together on one device.

The reference runs a case matrix as separate processes, packed onto a GPU by their memory
needs (``scripts/experiments/_runner_common.py:87-106``).  Here every case is an instance of one
context: a step is one launch (or one window + one bulk launch) over all of them.  The bulk
rows of all instances form one row list, cut by per-instance tile tables (a binary search
over the instances' first tiles maps a row to its instance); every marker window is one
workgroup.  ``Ensemble.autotune`` picks the launch schedule (workgroup size, rows in flight,
grid) by timing the candidates on the ensemble itself.

Instances of different sizes, collision models, immersed-boundary methods and delta kernels
may share a context; they must share the storage precision.
"""
from __future__ import annotations

import numpy as np

from .config import SimConfig
from .solver import Context, FixedCylinder

# Channel height (lattice nodes) of each Reynolds number in the 72-case matrix: the grids
# cycle through Ny = 641 .. 1921 with Re, Nx = 2.5 (Ny - 1) + 1 and D = (Ny - 1) / 16.
RE_NY = {20.0: 641, 40.0: 961, 100.0: 1281, 200.0: 1921}


def cylinder_cfg(Re, ny, ibm_method, delta_type, collision_model, lattice_u=0.04,
                 max_steps=200_000, check_interval=200):
    """Fixed cylinder of the Re = 100 family: channel 2.5 x 1 (Nx = 2.5 (Ny - 1) + 1),
    D = (Ny - 1) / 16 centred at (0.625, 0.5), one marker per lattice spacing, inflow U."""
    return SimConfig(Re=float(Re), NN=int(ny), xmax=2.5, ymax=1.0, cylinder_center=(0.625, 0.5),
                     cylinder_D_ratio=1 / 16, lattice_u=lattice_u, inflow_u=lattice_u,
                     bc_type="dirichlet_neumann", delta_type=delta_type, ibm_method=ibm_method,
                     collision_model=collision_model, trt_magic_param=0.25, mdf_iterations=20,
                     mdf_min_iterations=5, mdf_tolerance=1e-4, marker_spacing_factor=1.0,
                     max_steps=max_steps, check_interval=check_interval, use_convergence=False)


def case_matrix(res=(20.0, 40.0, 100.0, 200.0), methods=("DF", "MDF", "DFC"),
                deltas=("hat", "peskin4pt"), models=("BGK", "TRT", "CM_MRT"), shrink: int = 1,
                lattice_u=0.04):
    """The Re x method x kernel x collision matrix (4 x 3 x 2 x 3 = 72 cases by default).

    ``shrink`` divides (Ny - 1) by an integer, for small test versions of the same matrix."""
    cases = []
    for Re in res:
        ny = (RE_NY[float(Re)] - 1) // shrink + 1
        for m in methods:
            for d in deltas:
                for c in models:
                    cases.append(cylinder_cfg(Re, ny, m, d, c, lattice_u=lattice_u))
    return cases


class Ensemble:
    """Fixed-cylinder instances in one context, each with its own seeded perturbation.

    ``init="host"`` forms each instance's initial state as the reference does (numpy
    perturbation, pre-loop collision), ``"device"`` forms the same equilibrium on the device
    from a hash-seeded perturbation (for ensembles too large to set up on the host)."""

    def __init__(self, cases, precision="fp32dev", init="device", seed0=1000, amp=1e-3,
                 device=0):
        self.ctx = Context(device)
        self.members = [FixedCylinder(cfg, perturb=(seed0 + k, amp), ctx=self.ctx,
                                      precision=precision, init=init)
                        for k, cfg in enumerate(cases)]
        self.nodes = sum(m.st.nodenums for m in self.members)
        self.ttt = 0

    def advance(self, n, diag_every=0, graph=True):
        if n > 0:
            self.ctx.step(n, diag_every, graph)
            self.ttt += n
            for m in self.members:
                m.ttt += n

    def autotune(self, steps=10):
        return self.ctx.autotune(steps)

    def cd_cl(self):
        """(Cd, Cl) of every instance at the last diagnostic step."""
        return np.array([m.cd_cl() for m in self.members])

    def close(self):
        self.ctx.close()
