"""Host-side run control (no GPU): the time loop's termination bookkeeping, the settling stop
rules and the final step the reference reports.  Each case is worked out by hand from the
reference lines cited in the test."""
from types import SimpleNamespace

import numpy as np

from iblbm_amd import SimConfig
from iblbm_amd.sedimentation import stop_final_step, stop_reason
from iblbm_amd.solver import LoopControl


def _drive(cfg, fields):
    """Runs LoopControl the way solver.run does, with field k returned at check k."""
    ctl = LoopControl(cfg)
    ttt, k, hist = 0, 0, []
    while not ctl.done(ttt):
        ttt += cfg.check_interval
        hist.append(1.0)
        ctl.check(ttt, hist, lambda: fields[k])
        k += 1
    return ctl, ttt


def _field(v, n=12):
    U = np.zeros((n, 2))
    U[:, 0] = v
    return U


def test_convergence_at_last_check_ends_converged():
    # loop.py:64-70: the convergence exit is tested before the step bound
    cfg = SimConfig(Re=20.0, NN=21, max_steps=300, check_interval=100, use_convergence=True,
                    convergence_start=100, convergence_threshold=1e-6)
    ctl, ttt = _drive(cfg, [_field(1.0), _field(2.0), _field(2.0)])
    assert ttt == 300 and ctl.converged and ctl.termination == "converged"
    assert ctl.error == 0.0


def test_relative_l2_error_value():
    # core.py:116-137: sqrt(sum((u - u_old)^2) / sum(u_old^2)) over both components
    cfg = SimConfig(Re=20.0, NN=21, max_steps=200, check_interval=100, use_convergence=True,
                    convergence_start=100, convergence_threshold=1e-12)
    ctl, _ = _drive(cfg, [_field(2.0), _field(3.0)])
    assert not ctl.converged and ctl.termination == "max_steps"
    assert abs(ctl.error - 0.5) < 1e-15


def test_start_off_the_check_grid_never_converges():
    # loop.py:120-130: the reference field is taken only when ttt == convergence_start
    cfg = SimConfig(Re=20.0, NN=21, max_steps=500, check_interval=100, use_convergence=True,
                    convergence_start=150, convergence_threshold=1.0)
    ctl, ttt = _drive(cfg, [_field(1.0)] * 5)
    assert ttt == 500 and not ctl.converged and ctl.termination == "max_steps"
    assert ctl.pre_U is None and ctl.error == 1.0


def test_converged_before_the_end_stops_there():
    cfg = SimConfig(Re=20.0, NN=21, max_steps=1000, check_interval=100, use_convergence=True,
                    convergence_start=200, convergence_threshold=1e-3)
    ctl, ttt = _drive(cfg, [_field(5.0), _field(1.0), _field(1.0), _field(9.0)])
    # checks at 100 (before the start), 200 (reference field), 300 (error 0 -> converged)
    assert ttt == 300 and ctl.termination == "converged"


def test_cd_window_criterion():
    cfg = SimConfig(Re=20.0, NN=21, max_steps=10_000, check_interval=100, use_convergence=False,
                    cd_convergence_tol=1e-3, cd_convergence_start_step=250,
                    cd_convergence_window=3)
    ctl, ttt = _drive(cfg, [_field(1.0)] * 100)
    # the window is full at 300 (> 250): spread 0 < tol
    assert ttt == 300 and ctl.termination == "cd_converged" and ctl.converged


def test_final_step_of_a_stopped_run():
    # loop.py:106-109 breaks before ttt += 1; final_step = ttt - 1 (:142-148)
    assert stop_final_step(4200, "contact") == 4199
    assert stop_final_step(4200, "domain_bounds") == 4199
    assert stop_final_step(5000, "max_steps") == 5000
    assert stop_final_step(5000, None) == 5000


def _p(x, y, r=0.05, vx=0.0, vy=-0.01):
    return SimpleNamespace(pos=np.array([x, y]), vel=np.array([vx, vy]), r_domain=r)


def _scfg(**kw):
    base = dict(xmax=1.0, ymax=8.0, cylinder_D_ratio=0.1, gravity_direction="down",
                sedimentation_stop_at_contact=False, sedimentation_stop_offset_d=0.0)
    base.update(kw)
    return SimpleNamespace(**base)


def test_stop_rules_single_particle():
    dx = 1.0 / 256
    cfg = _scfg()
    assert stop_reason(cfg, [_p(0.5, 4.0)], dx, False, vy_star=1.0) is None
    # nan: the record's vy_star (sedimentation_single.py:343-344)
    assert stop_reason(cfg, [_p(0.5, 4.0)], dx, False, vy_star=float("nan")) == "nan"
    # contact at y <= r (sedimentation_common.py:20-26)
    c = _scfg(sedimentation_stop_at_contact=True)
    assert stop_reason(c, [_p(0.5, 0.05)], dx, False, vy_star=1.0) == "contact"
    assert stop_reason(c, [_p(0.5, 0.0501)], dx, False, vy_star=1.0) is None
    # offset at y <= offset_d * D (:27-33)
    o = _scfg(sedimentation_stop_offset_d=2.0)
    assert stop_reason(o, [_p(0.5, 0.2)], dx, False, vy_star=1.0) == "offset"
    assert stop_reason(o, [_p(0.5, 0.2001)], dx, False, vy_star=1.0) is None
    # domain bounds within 2 dx of the left wall (:37-66)
    assert stop_reason(cfg, [_p(0.05 + 1.5 * dx, 4.0)], dx, False, vy_star=1.0) == \
        "domain_bounds"
    # the bottom wall counts only when no terminal rule is set
    assert stop_reason(cfg, [_p(0.5, 0.05 + dx)], dx, False, vy_star=1.0) == "domain_bounds"
    assert stop_reason(c, [_p(0.5, 0.05 + dx)], dx, False, vy_star=1.0) is None
    # gravity "right": contact at x >= xmax - r, and the right wall is the terminal one
    r = _scfg(gravity_direction="right", sedimentation_stop_at_contact=True)
    assert stop_reason(r, [_p(0.95, 4.0)], dx, False, vy_star=1.0) == "contact"
    assert stop_reason(r, [_p(0.94, 4.0)], dx, False, vy_star=1.0) is None


def test_stop_rules_several_particles():
    dx = 1.0 / 192
    cfg = _scfg()
    a, b = _p(0.5, 4.0), _p(0.5, 5.0)
    assert stop_reason(cfg, [a, b], dx, True) is None
    # nan on particle 0's velocity along gravity (sedimentation_multi.py:302-305)
    assert stop_reason(cfg, [_p(0.5, 4.0, vy=float("inf")), b], dx, True) == "nan"
    assert stop_reason(cfg, [a, _p(0.5, 5.0, vy=float("nan"))], dx, True) is None
    g = _scfg(gravity_direction="right")
    assert stop_reason(g, [_p(0.5, 4.0, vx=float("nan")), b], dx, True) == "nan"
    # every particle is checked against the walls (:309-312)
    assert stop_reason(cfg, [a, _p(0.99, 5.0)], dx, True) == "domain_bounds"
    # but only particle 0 against the stopping rules
    c = _scfg(sedimentation_stop_at_contact=True)
    assert stop_reason(c, [a, _p(0.5, 0.04)], dx, True) is None
    assert stop_reason(c, [_p(0.5, 0.04), b], dx, True) == "contact"


def test_stream_first_runs_equal_the_default_order():
    """The reference's runs with stream_first=True (a collision before the loop, then Stream
    -> Macro -> IBM -> Collision each step) are its default-order runs bit for bit, for MDF
    and for DFC (fixtures written from the reference): the same operations in the same
    sequence, the loop boundary moved.  The device runs stream_first unchanged."""
    from conftest import golden, golden_cfg
    from iblbm_amd.init import initialize
    from iblbm_amd.solver import make_case
    for base, sf in (("run_mdf_trt_p4", "run_sf_mdf_trt_p4"),
                     ("run_dfc_cm_hat", "run_sf_dfc_cm_hat")):
        a, sa = golden(base)
        b, sb = golden(sf)
        assert sb["cfg"]["stream_first"] is True and not sa["cfg"].get("stream_first", False)
        for k in a.files:
            assert np.array_equal(a[k], b[k]), (sf, k)
        cfg = golden_cfg(sb)
        make_case(cfg, initialize(cfg, fields=False))  # accepted
