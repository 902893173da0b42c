"""Batched instances: a mixed ensemble in one context (one step launch for all instances)
against the reference's own runs, bit for bit, under several launch schedules and after the
autotuner (which runs trial steps and must restore the state exactly)."""
import numpy as np
import pytest

from conftest import RUN_CASES, golden, golden_cfg, perturb_of
from iblbm_amd.init import feq_host
from iblbm_amd.solver import Context, FixedCylinder

pytestmark = pytest.mark.gpu



def _fused_alone(case):
    g, sc = golden(case)
    sim = FixedCylinder(golden_cfg(sc), perturb=perturb_of(sc))
    mode = sim.ctx.step_mode()
    sim.ctx.close()
    return mode == "fused"


def _run_ensemble(cases, schedule):
    ctx = Context()
    members = []
    for c in cases:
        g, sc = golden(c)
        members.append((c, FixedCylinder(golden_cfg(sc), perturb=perturb_of(sc), ctx=ctx), g,
                        sc["final_step"]))
    if schedule == "autotune":
        tuned = ctx.autotune(4)
        print("autotuned:", tuned)
    elif schedule is not None:
        ctx.set_schedule(**schedule)
    mode = ctx.step_mode()
    t = 0
    for target in sorted({m[3] for m in members}):
        ctx.step(target - t, 0, graph=True)
        t = target
        for name, sim, g, final in members:
            if final != target:
                continue
            np.testing.assert_array_equal(sim.ctx.download_f(sim.k), g["f_post"], err_msg=name)
            np.testing.assert_array_equal(sim.ctx.download_fib(sim.k), g["fib"], err_msg=name)
    ctx.close()
    return mode


@pytest.mark.parametrize("schedule", [
    None,
    dict(mode="two_streams", bulk_threads=1024, bulk_rows=2, bulk_waves=0),
    dict(mode="two_streams", bulk_threads=512, bulk_rows=1, bulk_waves=40),
    dict(mode="serial", bulk_threads=64, bulk_rows=2, bulk_waves=7),
    dict(mode="two_streams", bulk_threads=512, bulk_rows=1, bulk_waves=0, stores="nontemporal"),
    dict(mode="serial", bulk_threads=256, bulk_rows=2, bulk_waves=0, stores="nontemporal"),
    "autotune",
])
def test_mixed_ensemble_matches_reference_runs(schedule):
    # BGK/TRT/CM-MRT x DF/MDF/DFC x hat/P4 on two grid sizes, one context
    mode = _run_ensemble(RUN_CASES, schedule)
    assert mode in ("two_streams", "serial")


@pytest.mark.parametrize("schedule", [None, dict(step_rows=1), dict(step_order="blocked"),
                                      dict(step_rows=1, step_order="blocked"),
                                      dict(step_order=7), dict(step_order=8, bulk_delay_us=2.0),
                                      dict(step_rows=1, step_order="queue"),
                                      dict(step_rows=2, step_order="queue", bulk_delay_us=1.0),
                                      dict(step_rows=3, step_order="queue"),
                                      dict(step_rows=3, step_order="queue", bulk_delay_us=1.0),
                                      "autotune"])
def test_fused_ensemble_matches_reference_runs(schedule):
    # the cases whose window qualifies for the fused launch: one kernel per step for all
    fast = [c for c in RUN_CASES if _fused_alone(c)]
    print("fused-eligible:", fast)
    assert len(fast) >= 3
    mode = _run_ensemble(fast, schedule)
    if schedule != "autotune":
        assert mode == "fused"


def test_device_initial_state():
    g, sc = golden("run_df_bgk_hat")
    cfg = golden_cfg(sc)
    ctx = Context()
    a = FixedCylinder(cfg, ctx=ctx, init="device")
    b = FixedCylinder(cfg, perturb=(7, 1e-3), ctx=ctx, init="device")
    c = FixedCylinder(cfg, perturb=(7, 1e-3), ctx=ctx, init="device")
    d = FixedCylinder(cfg, perturb=(8, 1e-3), ctx=ctx, init="device")
    N = a.st.nodenums
    # no perturbation: the uniform-inflow equilibrium, the same arithmetic as the host's
    U = np.zeros((N, 2))
    U[:, 0] = cfg.inflow_u
    np.testing.assert_array_equal(ctx.download_f(a.k), feq_host(np.ones(N), U))
    fb = ctx.download_f(b.k)
    np.testing.assert_array_equal(fb, ctx.download_f(c.k))  # same seed, same field
    assert not np.array_equal(fb, ctx.download_f(d.k))
    rho = fb.sum(axis=1)
    ex = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1], dtype=float)
    ux = (fb * ex).sum(axis=1) / rho
    xi = ux / cfg.inflow_u - 1.0
    assert np.max(np.abs(rho - 1.0)) < 1e-14
    assert np.max(np.abs(xi)) <= 1e-3 * (1 + 1e-9)
    assert abs(np.mean(xi)) < 1e-4 and abs(np.std(xi) - 1e-3 / np.sqrt(3)) < 1e-4
    ctx.close()


@pytest.mark.parametrize("schedule", [dict(step_rows=3), dict(step_rows=4),
                                      dict(mode="serial", bulk_rows=3),
                                      dict(mode="two_streams", bulk_rows=4, bulk_threads=512)])
def test_fp64_takes_two_rows_for_the_fp32_only_schedules(schedule):
    """Rows-in-flight values 3 and 4 apply to FP32 storage; an FP64 context runs them as two
    rows in flight, with unchanged results."""
    fast = [c for c in RUN_CASES if _fused_alone(c)][:2]
    _run_ensemble(fast, schedule)
