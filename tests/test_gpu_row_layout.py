"""Bulk row layouts: 64-column tiles of lattice rows, or 64 consecutive row-major nodes per
bulk row ("flat", for FP32-deviation contexts whose instances are at least 64 nodes wide;
FP64 contexts keep tiles).  The layout only changes which wavefront updates which node, so
every result is bit-identical: the FP32 populations of both layouts are compared bit for bit
on awkward widths (chunks that end on, before and after a row end) for the DF, MDF and DFC
windows in the fused and two-kernel launches, and FP64 contexts asked for the flat layout
still match the oracle and the reference's settling run."""
import numpy as np
import pytest

from conftest import golden, golden_cfg, rel_linf
from iblbm_amd import SimConfig
from iblbm_amd.solver import Context, FixedCylinder
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _cyl(nn, xmax, method, delta, coll, steps):
    return SimConfig(Re=10.0, NN=nn, xmax=xmax, ymax=1.0, cylinder_center=(0.3 * xmax, 0.5),
                     cylinder_D_ratio=0.25, lattice_u=0.05, inflow_u=0.05,
                     bc_type="dirichlet_neumann", delta_type=delta, ibm_method=method,
                     collision_model=coll, marker_spacing_factor=1.0, mdf_iterations=20,
                     mdf_min_iterations=5, mdf_tolerance=1e-4, max_steps=steps,
                     check_interval=steps, use_convergence=False)


def _run(cfg, steps, layout, precision="fp64", mode="auto", expect=None):
    ctx = Context()
    ctx.set_schedule(mode=mode, row_layout=layout)
    sim = FixedCylinder(cfg, ctx=ctx, precision=precision)
    assert sim.ctx.schedule()["row_layout"] == layout
    if expect is not None:
        assert sim.ctx.row_layout_in_use() == expect
    sim.advance(steps, diag_every=steps)
    f = sim.ctx.download_f(sim.k)
    cd = sim.cd_cl()[0]
    ctx.close()
    return f, cd


@pytest.mark.parametrize("layout", ["tiles", "flat"])
@pytest.mark.parametrize("case", [("DF", "hat", "BGK"), ("MDF", "peskin4pt", "TRT"),
                                  ("DFC", "hat", "CM_MRT")])
def test_layout_matches_oracle_on_a_wide_channel(case, layout):
    steps = 60
    cfg = _cyl(25, 5.0, *case, steps)  # 121 x 25 nodes
    print(f"row layout {layout}: 121x25 channel {case}, {steps} steps vs oracle", flush=True)
    f, cd = _run(cfg, steps, layout)
    ref = O.FixedCylinder(cfg).run(steps)
    np.testing.assert_array_equal(f, ref.f_post)
    assert abs(cd - ref.Cd[-1]) <= 1e-10 * abs(ref.Cd[-1])


@pytest.mark.parametrize("case", [("DF", "hat", "BGK"), ("MDF", "peskin4pt", "TRT"),
                                  ("DFC", "hat", "CM_MRT")])
@pytest.mark.parametrize("nn,xmax", [(17, 4.0), (17, 8.0), (33, 2.0), (17, 8.0625),
                                     (25, 5.0)])
@pytest.mark.parametrize("mode", ["fused", "two_streams"])
def test_layouts_agree_in_fp32(nn, xmax, mode, case):
    """Widths 65, 129, 65, 130 and 121 nodes."""
    steps = 40
    cfg = _cyl(nn, xmax, *case, steps)
    print(f"row layouts in fp32: NN={nn} xmax={xmax} {mode} {case}", flush=True)
    f_t, cd_t = _run(cfg, steps, "tiles", "fp32dev", mode)
    f_f, cd_f = _run(cfg, steps, "flat", "fp32dev", mode, expect="flat")
    np.testing.assert_array_equal(f_t, f_f)
    assert abs(cd_t - cd_f) <= 1e-12 * abs(cd_t)


def test_flat_layout_settling_matches_reference():
    from iblbm_amd.sedimentation import Sedimentation
    g, sc = golden("sed_single_mdf_cm_p4")  # 129 x 193, settling channel
    cfg = golden_cfg(sc)
    ctx = Context()
    ctx.set_schedule(row_layout="flat")
    print("row layout flat: settling disc 129x193, 40 steps vs reference", flush=True)
    sim = Sedimentation(cfg, ctx=ctx)
    for _ in range(sc["final_step"]):
        sim.step()
    p = sim.particles[0]
    row = np.array([p.pos[0], p.pos[1], p.vel[0], p.vel[1]])
    assert rel_linf(row, g["particles"][-1, 1:5]) <= 1e-10
    np.testing.assert_array_equal(sim.ctx.download_f(sim.k), g["f_post"])
    ctx.close()


@pytest.mark.parametrize("stores", ["normal", "nontemporal"])
@pytest.mark.parametrize("layout", ["tiles", "flat"])
def test_store_modes_and_layouts_agree_on_settling_members(layout, stores):
    """Two full-size two-disc members (193 x 1921, FP32 deviations): every layout and store
    mode of the serial two-kernel launch gives the populations of the default schedule."""
    from iblbm_amd.sed_ensemble import SedimentationEnsemble, two_disc_cfgs
    cfgs = two_disc_cfgs(2, seed=9, check_interval=50)
    out = []
    for sched in (None, dict(mode="serial", bulk_threads=256, row_layout=layout,
                             stores=stores)):
        ens = SedimentationEnsemble(cfgs, precision="fp32dev")
        if sched is not None:
            ens.ctx.set_schedule(**sched)
            assert ens.ctx.row_layout_in_use() == layout
        print(f"settling members, schedule {sched}, 50 steps", flush=True)
        ens.run(50)
        out.append([ens.ctx.download_f(m.k) for m in ens.members])
        ens.close()
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)
