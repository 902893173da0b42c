"""run(cfg, save_dir=...) on the GPU writes the reference's output files, with the run's own
numbers: the fixed-cylinder Cd history and the settling trajectory."""
import dataclasses
import json

import numpy as np
import pytest

from conftest import golden, golden_cfg

pytestmark = pytest.mark.gpu


def test_fixed_cylinder_run_writes_status(tmp_path):
    from iblbm_amd.solver import run
    g, sc = golden("run_df_bgk_hat")
    cfg = golden_cfg(sc)
    print("fixed cylinder run with save_dir: 41x21 BGK/DF/hat", flush=True)
    res = run(cfg, verbose=False, save_dir=str(tmp_path))
    st = json.load(open(tmp_path / "status.json"))
    assert st["completed"] and st["final_step"] == res["final_step"]
    assert np.array_equal(np.array(st["Cd_history"]), res["Cd_history"])
    assert st["termination_reason"] == res["termination_reason"]


def test_settling_run_writes_history(tmp_path):
    from iblbm_amd.sedimentation import run
    g, sc = golden("sed_multi_dfc_cm_hat")
    cfg = dataclasses.replace(golden_cfg(sc), check_interval=10, max_steps=60)
    print("settling run with save_dir: two discs DFC 25x97, 60 steps", flush=True)
    res = run(cfg, verbose=False, save_dir=str(tmp_path))
    h = json.load(open(tmp_path / "sedimentation_history.json"))
    st = json.load(open(tmp_path / "status.json"))
    assert len(h) == 6 and st["completed"] and st["final_step"] == 60
    ref = g["particles"]  # rows: step, then 9 columns per particle
    for rec in h:
        k = rec["step"] - 1
        for pi in range(2):
            assert rec["particles"][pi]["y"] == ref[k, 1 + 9 * pi + 1]
            assert rec["particles"][pi]["vy"] == ref[k, 1 + 9 * pi + 3]
    assert len(st["Re_standard_peaks"]) == 2
