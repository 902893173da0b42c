"""The learned MDF force predictor on the device: FP32 gradients against the CPU oracle, BF16
MFMA inference against its emulation, and what the trained network gains over plain MDF and
over a warm start without the network."""
import numpy as np
import pytest

from iblbm_amd import SimConfig
from iblbm_amd.predictor import mlp_infer_bf16, mlp_loss_grad
from iblbm_amd.solver import FixedCylinder
from oracle import mlp_ref as M

pytestmark = pytest.mark.gpu


def _glorot(seed):
    rng = np.random.default_rng(seed)
    w = {}
    for name, shape in M.SHAPES:
        if name.startswith("W"):
            a = np.sqrt(6.0 / (shape[0] + shape[1]))
            w[name] = rng.uniform(-a, a, shape)
        else:
            w[name] = rng.standard_normal(shape) * 0.1
    return M.pack(w).astype(np.float32).astype(np.float64)


@pytest.mark.parametrize("rows", [128, 37])
def test_device_gradient_matches_oracle(rows):
    p = _glorot(3)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((rows, M.IN)).astype(np.float32).astype(np.float64)
    y = rng.standard_normal((rows, M.OUT)).astype(np.float32).astype(np.float64)
    loss, g = mlp_loss_grad(p, x, y)
    lo, go = M.loss_grad(p, x, y)
    assert abs(loss - lo) <= 1e-5 * abs(lo)
    # FP32 device arithmetic against the FP64 oracle
    err = np.max(np.abs(g - go)) / np.max(np.abs(go))
    print(f"rows={rows}: loss {loss:.6f} vs {lo:.6f}, gradient rel Linf {err:.2e}")
    assert err <= 1e-4


def test_bf16_mfma_inference_matches_emulation():
    p = _glorot(5)
    rng = np.random.default_rng(6)
    x = rng.standard_normal((200, M.IN)).astype(np.float32)
    y = mlp_infer_bf16(p, x)
    ye = M.forward_bf16(p, x)
    y64 = M.forward(p, x)[0]
    scale = np.max(np.abs(y64))
    e_emu = np.max(np.abs(y - ye)) / scale
    e_64 = np.max(np.abs(y - y64)) / scale
    print(f"BF16 inference: vs emulation {e_emu:.2e}, vs FP64 {e_64:.2e}")
    assert e_emu <= 2e-2 and e_64 <= 5e-2


def _cfg():
    return SimConfig(Re=100.0, NN=129, xmax=2.5, ymax=1.0, cylinder_center=(0.5, 0.5),
                     cylinder_D_ratio=0.125, lattice_u=0.04, inflow_u=0.04,
                     bc_type="dirichlet_neumann", delta_type="peskin4pt", ibm_method="MDF",
                     collision_model="TRT", trt_magic_param=0.25, mdf_iterations=20,
                     mdf_min_iterations=5, mdf_tolerance=1e-4, marker_spacing_factor=1.0,
                     max_steps=10_000, check_interval=100, use_convergence=False)


def _held_out_mse(pred, advance, steps):
    """Network and warm-start errors on `steps` further plain-MDF steps (not trained on).

    Rows hold the scaled features and the scaled change of the converged force; the warm
    start predicts no change, so its error is the change itself.  Both are normalised by the
    training set's target scale."""
    n0 = pred.samples_written()
    advance(steps)
    rows = pred.data(n0, pred.samples_written() - n0).astype(np.float64)
    p, nrm = pred.params()
    x = (rows[:, :M.IN] - nrm[:M.IN]) / nrm[M.IN:2 * M.IN]
    y = rows[:, M.IN:] / nrm[2 * M.IN:]
    return float(np.mean((M.forward(p, x)[0] - y) ** 2)), float(np.mean(y ** 2))


def test_learned_predictor_gain():
    sim = FixedCylinder(_cfg(), perturb=(7, 1e-3))
    assert sim.ctx.step_mode() == "fused"
    pred = sim.predictor
    sim.advance(400)  # past the start-up transient
    pred.mode("collect", capacity=1 << 18)
    sim.advance(2000)
    n = pred.samples()
    assert n >= 1000 * sim.st.Lb
    losses = pred.train(steps=3000, lr=2e-3, seed=11)
    print(f"samples {n}, loss first/last 100: {losses[:100].mean():.4e} {losses[-100:].mean():.4e}")
    assert losses[-100:].mean() < 0.25 * losses[:100].mean()
    # out-of-sample accuracy of the predicted change of force, on the next 100 steps
    mse_net, mse_warm = _held_out_mse(pred, sim.advance, 100)
    print(f"held-out force change error (normalised MSE): network {mse_net:.3e}, "
          f"warm start {mse_warm:.3e}")
    assert mse_net < 0.1 * mse_warm

    # every mode restarts from the same populations and previous forces; "collect" is plain
    # MDF that also counts its sub-steps
    f0 = sim.ctx.download_f(sim.k)
    F0 = pred.previous_force(sim.st.Lb)
    res = {}
    for mode in ("collect", "warm", "learned"):
        sim.ctx.upload_f(sim.k, f0)
        pred.set_previous_force(F0)
        pred.mode(mode)
        pred.stats(reset=True)
        sim.advance(300, diag_every=300)
        res[mode] = (sim.cd_cl()[0], pred.stats())
        print(mode, res[mode])
    cd_plain = res["collect"][0]
    for mode in ("warm", "learned"):
        assert abs(res[mode][0] - cd_plain) <= 1e-3 * abs(cd_plain)
        assert res[mode][1]["restarts"] == 0
    # fewer true MDF sub-steps than plain MDF (which always takes at least 5)
    assert res["learned"][1]["substeps_per_step"] < 0.5 * res["collect"][1]["substeps_per_step"]


def _sed_cfg():
    return SimConfig(Re=38.4, NN=129, xmax=1.0, ymax=1.5, cylinder_center=(0.5, 1.0),
                     cylinder_D_ratio=0.25, rho_ratio=1.5, gravity=2.7e-5,
                     motion_type="sedimentation", bc_type="settling_channel",
                     collision_model="CM_MRT", ibm_method="MDF", delta_type="peskin4pt",
                     mdf_iterations=20, mdf_min_iterations=5, mdf_tolerance=1e-4,
                     marker_spacing_factor=1.0, settling_inertia_model="explicit_history",
                     enable_rotation=True, max_steps=100_000, check_interval=100)


def test_learned_predictor_on_a_settling_disc():
    from iblbm_amd.sedimentation import Sedimentation
    sim = Sedimentation(_sed_cfg())
    assert sim.ctx.step_mode() == "fused"
    pred = sim.predictor
    for _ in range(200):
        sim.step()
    pred.mode("collect", capacity=1 << 18)
    for _ in range(600):
        sim.step()
    losses = pred.train(steps=3000, lr=2e-3, seed=3)
    print(f"settling: samples {pred.samples()}, loss first/last 100 "
          f"{losses[:100].mean():.3e} {losses[-100:].mean():.3e}")

    def advance(n):
        for _ in range(n):
            sim.step()

    mse_net, mse_warm = _held_out_mse(pred, advance, 100)
    print(f"settling held-out force change error: network {mse_net:.3e}, warm {mse_warm:.3e}")
    assert mse_net < 0.1 * mse_warm
    snap = sim.snapshot()
    F0 = pred.previous_force(sim.particles[0].Lb)
    res = {}
    for mode in ("collect", "warm", "learned"):
        sim.restore(snap)
        pred.set_previous_force(F0)
        pred.mode(mode)
        pred.stats(reset=True)
        for _ in range(200):
            sim.step()
        p0 = sim.particles[0]
        res[mode] = (float(p0.vel[1]), float(p0.pos[1]), pred.stats())
        print(mode, res[mode])
    v_plain, y_plain = res["collect"][0], res["collect"][1]
    for mode in ("warm", "learned"):
        assert abs(res[mode][0] - v_plain) <= 1e-3 * abs(v_plain)
        assert res[mode][2]["restarts"] == 0
    assert res["learned"][2]["substeps_per_step"] < 0.5 * res["collect"][2]["substeps_per_step"]
