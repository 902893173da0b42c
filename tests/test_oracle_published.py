"""The oracle against the reference's published Taylor-Green convergence table
(``data/taylor_green/df/convergence.npz``, ``data/taylor_green/{df,dfc}/hat/convergence.npz``,
copied into tests/golden/published.npz): the pure-lattice L2 error at D = 10 and 20 to 1e-11
relative, the hat-kernel DF and DFC errors at D = 10 to 2e-6 (the table's own state of the
code), which pins the analytic boundary, the initial field and the markers' desired velocity
of the oracle to published numbers as well as to the fixtures."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
from test_gpu_published import PUB, tg_config  # noqa: E402  (configuration builder only)
from iblbm_amd import taylor_green as TG  # noqa: E402


def _oracle_error(cfg):
    import oracle.oracle as O
    sim = O.FixedCylinder(cfg)
    for _ in range(cfg.max_steps):
        sim.step()
    return TG.l2_error(sim.st.U, TG.analytic_velocity(cfg, sim.st.nx, sim.st.ny,
                                                      cfg.max_steps * sim.st.dt))


@pytest.mark.parametrize("name,method,delta,D,tol", [
    ("lbm_p4", None, "peskin4pt", 10, 1e-11), ("lbm_p4", None, "peskin4pt", 20, 1e-11),
    ("df_hat", "DF", "hat", 10, 2e-6), ("dfc_hat", "DFC", "hat", 10, 2e-6)])
def test_oracle_reproduces_published_taylor_green(name, method, delta, D, tol):
    k = list(PUB[f"tg_{name}_resolutions"]).index(D)
    e = _oracle_error(tg_config(D, method, delta))
    assert abs(e / PUB[f"tg_{name}_errors"][k] - 1.0) <= tol


def test_planted_boundary_error_is_caught(monkeypatch):
    """A wrong decay time in the analytic boundary (t - 1 instead of t) moves the pure-lattice
    error at D = 10 far outside the bar above."""
    import oracle.oracle as O
    real = O.tg_boundary_velocity
    monkeypatch.setattr(O, "tg_boundary_velocity", lambda cfg, nx, ny, t: real(cfg, nx, ny, t - 1))
    e = _oracle_error(tg_config(10))
    assert abs(e / PUB["tg_lbm_p4_errors"][0] - 1.0) > 1e-6
