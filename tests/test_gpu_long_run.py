"""1,000-step FP64 runs against the oracle: populations within 1e-12 relative L-infinity (the
measured error is printed; the kernels reproduce the reference arithmetic, so it is 0), and
Cd / Cl at every check interval within 1e-10 relative, for BGK + DF + hat, TRT + MDF +
Peskin-4 and CM-MRT + DFC + hat on a 121 x 25 channel with a seeded 1e-3 perturbation of
the initial velocity."""
import numpy as np
import pytest

from conftest import rel_linf
from iblbm_amd import SimConfig
from iblbm_amd.solver import FixedCylinder
from oracle import oracle as O

pytestmark = pytest.mark.gpu
STEPS = 1000


def _cfg(method, delta, coll):
    return SimConfig(Re=10.0, NN=25, xmax=5.0, ymax=1.0, cylinder_center=(1.5, 0.5),
                     cylinder_D_ratio=0.25, lattice_u=0.05, inflow_u=0.05,
                     bc_type="dirichlet_neumann", delta_type=delta, ibm_method=method,
                     collision_model=coll, marker_spacing_factor=1.0, mdf_iterations=20,
                     mdf_min_iterations=5, mdf_tolerance=1e-4, trt_magic_param=0.25,
                     max_steps=STEPS, check_interval=100, use_convergence=False)


@pytest.mark.parametrize("case", [("DF", "hat", "BGK"), ("MDF", "peskin4pt", "TRT"),
                                  ("DFC", "hat", "CM_MRT")])
def test_thousand_steps_match_oracle(case):
    cfg = _cfg(*case)
    perturb = (77, 1e-3)
    print(f"1000-step fp64 run vs oracle: 121x25 {case}", flush=True)
    sim = FixedCylinder(cfg, perturb=perturb)
    cd, cl = [], []
    for _ in range(STEPS // cfg.check_interval):
        sim.advance(cfg.check_interval, diag_every=cfg.check_interval, graph=True)
        c, l = sim.cd_cl()
        cd.append(c)
        cl.append(l)
    f = sim.ctx.download_f(sim.k)
    sim.ctx.close()
    ref = O.FixedCylinder(cfg, perturb=perturb).run(STEPS)
    err = rel_linf(f, ref.f_post)
    print(f"{case}: populations rel Linf {err:.3e} (bit-identical: "
          f"{np.array_equal(f, ref.f_post)}), Cd {cd[-1]:.10f}", flush=True)
    assert err <= 1e-12
    assert rel_linf(cd, ref.Cd) <= 1e-10
    assert np.max(np.abs(np.array(cl) - np.array(ref.Cl)) / np.abs(np.array(ref.Cd))) <= 1e-10


def _c1_cfg(steps):
    """401 x 161 channel, D = 20 at (100, 80), 63 markers, U = 0.05, Re = 20: BGK + DF + hat,
    Dirichlet inflow / Neumann outflow."""
    return SimConfig(Re=20.0, NN=161, xmax=2.5, ymax=1.0, cylinder_center=(0.625, 0.5),
                     cylinder_D_ratio=0.125, lattice_u=0.05, inflow_u=0.05,
                     bc_type="dirichlet_neumann", delta_type="hat", ibm_method="DF",
                     collision_model="BGK", marker_spacing_factor=1.0, mdf_iterations=1,
                     max_steps=steps, check_interval=100, use_convergence=False)


def _c2_cfg(steps):
    """1601 x 641 channel, D = 40, 126 markers, U = 0.04, Re = 100: TRT (Lambda = 1/4) +
    MDF [5, 20] tol 1e-4 + Peskin-4 (the benchmark workload)."""
    return SimConfig(Re=100.0, NN=641, xmax=2.5, ymax=1.0, cylinder_center=(0.625, 0.5),
                     cylinder_D_ratio=1.0 / 16.0, lattice_u=0.04, inflow_u=0.04,
                     bc_type="dirichlet_neumann", delta_type="peskin4pt", ibm_method="MDF",
                     collision_model="TRT", trt_magic_param=0.25, marker_spacing_factor=1.0,
                     mdf_iterations=20, mdf_min_iterations=5, mdf_tolerance=1e-4,
                     max_steps=steps, check_interval=100, use_convergence=False)


@pytest.mark.parametrize("which,steps,perturb", [("C1", 1000, None), ("C2", 1000, (1234, 1e-3))])
def test_benchmark_configs_match_oracle(which, steps, perturb):
    """The two fixed-cylinder configurations the benchmark is quoted on, at full size, for
    1,000 steps against the oracle: populations within 1e-12 relative L-infinity (bit for bit
    in practice), Cd at every check interval within 1e-10 relative."""
    O.set_threads(16)
    cfg = _c1_cfg(steps) if which == "C1" else _c2_cfg(steps)
    print(f"{which} at full size vs oracle, {steps} steps", flush=True)
    sim = FixedCylinder(cfg, perturb=perturb)
    assert (sim.st.nx, sim.st.ny) == ((401, 161) if which == "C1" else (1601, 641))
    cd = []
    for _ in range(steps // cfg.check_interval):
        sim.advance(cfg.check_interval, diag_every=cfg.check_interval, graph=True)
        cd.append(sim.cd_cl()[0])
    f = sim.ctx.download_f(sim.k)
    sim.ctx.close()
    ref = O.FixedCylinder(cfg, perturb=perturb).run(steps)
    err = rel_linf(f, ref.f_post)
    print(f"{which}: populations rel Linf {err:.3e} (bit-identical: "
          f"{np.array_equal(f, ref.f_post)}), Cd {cd[-1]:.10f}", flush=True)
    assert err <= 1e-12
    assert rel_linf(cd, ref.Cd) <= 1e-10


@pytest.mark.parametrize("precision", ["fp64", "fp32dev"])
def test_benchmark_workload_schedules_bit_identical(precision):
    """The benchmark workload at full size (1601 x 641, TRT + MDF + Peskin-4), 400 steps under
    the fused launch's row orders: the workgroup row queue taking one or two rows, and one row
    with the next row's populations copied into LDS ahead (with and without the bulk start
    delay) give the default schedule's populations and Cd bit for bit."""
    cfg = _c2_cfg(400)
    scheds = [None, dict(step_rows=1, step_order="queue"), dict(step_rows=2, step_order="queue"),
              dict(step_rows=3, step_order="queue"),
              dict(step_rows=3, step_order="queue", bulk_delay_us=1.0)]
    outs = []
    for sc in scheds:
        print(f"{precision} schedule {sc}", flush=True)
        sim = FixedCylinder(cfg, perturb=(1234, 1e-3), precision=precision)
        if sc is not None:
            sim.ctx.set_schedule(**sc)
            assert sim.ctx.step_mode() == "fused"
        cd = []
        for _ in range(4):
            sim.advance(100, diag_every=100, graph=True)
            cd.append(sim.cd_cl())
        outs.append((sim.ctx.download_f(sim.k), np.array(cd)))
        sim.ctx.close()
    for f, cd in outs[1:]:
        np.testing.assert_array_equal(f, outs[0][0])
        np.testing.assert_array_equal(cd, outs[0][1])

