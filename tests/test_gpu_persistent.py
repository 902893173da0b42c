"""The persistent step launch (``iblbm_step_persistent``): n time steps in one cooperative
launch of the fused step kernel with a grid-wide barrier between steps.  From the same state
it computes exactly what n per-step launches compute: populations, fib, MDF statistics and the
drag and lift of the last diagnostic step, in FP64 and FP32-deviation storage, for the three
collision models and with the learned predictor on; contexts it does not cover are refused.
"""
import dataclasses

import numpy as np
import pytest

from iblbm_amd import SimConfig
from iblbm_amd.solver import FixedCylinder

pytestmark = pytest.mark.gpu


def _cfg(model="TRT", ibm="MDF", delta="peskin4pt"):
    return SimConfig(Re=100.0, NN=129, xmax=2.5, ymax=1.0, cylinder_center=(0.5, 0.5),
                     cylinder_D_ratio=0.125, lattice_u=0.04, inflow_u=0.04,
                     bc_type="dirichlet_neumann", delta_type=delta, ibm_method=ibm,
                     collision_model=model, trt_magic_param=0.25, mdf_iterations=20,
                     mdf_min_iterations=5, mdf_tolerance=1e-4, marker_spacing_factor=1.0,
                     max_steps=10_000, check_interval=100, use_convergence=False)


def _state(sim):
    return (sim.ctx.download_f(sim.k), sim.ctx.download_fib(sim.k), sim.ctx.mdf_stats(sim.k),
            sim.cd_cl())


@pytest.mark.parametrize("prec,model", [("fp64", "TRT"), ("fp64", "BGK"), ("fp64", "CM_MRT"),
                                        ("fp32dev", "TRT"), ("fp32dev", "CM_MRT")])
def test_persistent_launch_equals_per_step_launches(prec, model):
    cfgs = _cfg(model)
    a = FixedCylinder(cfgs, perturb=(3, 1e-3), precision=prec)
    b = FixedCylinder(cfgs, perturb=(3, 1e-3), precision=prec)
    assert a.ctx.step_mode() == "fused" and b.ctx.step_mode() == "fused"
    for sim in (a, b):
        sim.advance(20)
    a.advance(150, diag_every=50)
    b.advance(150, diag_every=50, persistent=True)
    fa, fiba, sa, ca = _state(a)
    fb, fibb, sb, cb = _state(b)
    print(prec, model, "per-step", sa, ca, "persistent", sb, cb)
    np.testing.assert_array_equal(fb, fa)
    np.testing.assert_array_equal(fibb, fiba)
    assert sb == sa and cb == ca
    # a second persistent launch continues from the first (counter and parity carried over)
    a.advance(37, diag_every=37)
    b.advance(37, diag_every=37, persistent=True)
    np.testing.assert_array_equal(b.ctx.download_f(b.k), a.ctx.download_f(a.k))
    assert b.cd_cl() == a.cd_cl()
    a.ctx.close()
    b.ctx.close()


def test_persistent_launch_with_the_learned_predictor():
    sims = [FixedCylinder(_cfg(), perturb=(7, 1e-3)) for _ in range(2)]
    for sim in sims:
        sim.advance(200)
        pred = sim.predictor
        pred.mode("collect", capacity=1 << 16)
        sim.advance(200)
        pred.train(steps=200, lr=2e-3, seed=5)
        pred.mode("learned")
    sims[0].advance(60, diag_every=60)
    sims[1].advance(60, diag_every=60, persistent=True)
    np.testing.assert_array_equal(sims[1].ctx.download_f(sims[1].k),
                                  sims[0].ctx.download_f(sims[0].k))
    assert sims[1].predictor.stats() == sims[0].predictor.stats()
    assert sims[1].cd_cl() == sims[0].cd_cl()
    for sim in sims:
        sim.ctx.close()


def test_persistent_launch_refuses_what_it_does_not_cover():
    # a window above the fused path's 128 markers runs as two kernels: refused
    big = FixedCylinder(dataclasses.replace(_cfg(), NN=257, cylinder_D_ratio=0.25))
    assert big.ctx.step_mode() != "fused"
    with pytest.raises(ValueError):
        big.advance(5, persistent=True)
    big.ctx.close()
