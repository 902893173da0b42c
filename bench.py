"""Benchmark: MLUPS and HBM-roofline fraction of the fused IB-LBM step on one MI355X.

Workload (default): the Re=100 fixed cylinder on a 1601x641 lattice (D = 40, 126 markers),
TRT (Lambda = 1/4) + adaptive MDF (5..20 sub-steps, tol 1e-4) + Peskin 4-point kernel,
inflow U = 0.04 with a seeded 1e-3 velocity perturbation, FP64 populations.  A step is one
launch of the fused kernel (stream + boundary + moments + MDF window + collision) over the
whole lattice; the host reads the FP64 force reductions every check_interval (200) steps, as
the reference loop does.

Prints ONE JSON line.  ``--gpus N`` with N > 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
KERNEL_OF_MODE = {"fused": "k_step (window + bulk rows, one launch per step)",
                  "two_streams": "k_ibm || k_bulk (two streams, one step)",
                  "serial": "k_ibm then k_bulk (one stream, one step)"}
BYTES_PER_NODE = {"fp64": 144, "fp32dev": 72}  # 9 reads + 9 writes per node update


def workload_cfg(max_steps=200_000):
    from iblbm_amd import SimConfig
    return SimConfig(Re=100.0, NN=641, xmax=2.5, ymax=1.0, cylinder_center=(0.625, 0.5),
                     cylinder_D_ratio=1 / 16, lattice_u=0.04, inflow_u=0.04,
                     bc_type="dirichlet_neumann", delta_type="peskin4pt", ibm_method="MDF",
                     collision_model="TRT", trt_magic_param=0.25, mdf_iterations=20,
                     mdf_min_iterations=5, mdf_tolerance=1e-4, marker_spacing_factor=1.0,
                     max_steps=max_steps, check_interval=200, use_convergence=False)


PERTURB = (1234, 1e-3)


def cpu_info():
    cores = os.cpu_count() or 1
    model = platform.processor() or "unknown"
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except Exception:
        pass
    return model, cores


def cpu_baseline(seconds=15.0, threads=None):
    """The FP64 C/OpenMP oracle on the same workload: timed steps within a bounded budget."""
    from oracle import oracle as O
    model, cores = cpu_info()
    n = threads or min(cores, 16)
    used = O.set_threads(n)
    cfg = workload_cfg()
    sim = O.FixedCylinder(cfg, PERTURB)
    N = sim.st.N
    sim.step()  # warm-up
    t0 = time.perf_counter()
    steps = 0
    while time.perf_counter() - t0 < seconds:
        sim.step()
        steps += 1
    dt = time.perf_counter() - t0
    return {"value": N * steps / dt / 1e6, "unit": "MLUPS", "cores": used, "kind": "port",
            "sample": f"{steps} steps of the same 1601x641 TRT/MDF/P4 case after 1 warm-up "
                      f"step ({dt:.1f} s, {used} OpenMP threads, {model})"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--warmup", type=int, default=400)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--precision", choices=["fp64", "fp32dev"], default="fp64")
    ap.add_argument("--serial", action="store_true",
                    help="run the window and bulk kernels in order on one stream")
    ap.add_argument("--no-graph", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: the lattice fits one MI355X and this build has no "
              "multi-GPU path (see DESIGN.md for how it would shard)", file=sys.stderr)
        sys.exit(2)

    from iblbm_amd.solver import FixedCylinder
    cfg = workload_cfg()
    ci = cfg.check_interval
    steps = max(ci, (args.steps // ci) * ci)
    warm = max(ci, (args.warmup // ci) * ci)
    sim = FixedCylinder(cfg, perturb=PERTURB, precision=args.precision)
    if args.serial:
        sim.ctx.set_concurrent(False)
    graph = not args.no_graph
    mode = sim.ctx.step_mode()
    N = sim.st.nodenums
    for _ in range(warm // ci):
        sim.advance(ci, diag_every=ci, graph=graph)
        sim.cd_cl()
    sim.ctx.sync()
    ev_ms = 0.0
    iters = []
    t0 = time.perf_counter()
    for _ in range(steps // ci):
        sim.advance(ci, diag_every=ci, graph=graph)
        cd, cl = sim.cd_cl()  # synchronises, as the reference's diagnostics do
        ms, _n = sim.ctx.last_timing()
        ev_ms += ms
        iters.append(sim.ctx.mdf_stats(sim.k)["iter_count"])
    sim.ctx.sync()
    wall = time.perf_counter() - t0
    mlups = N * steps / wall / 1e6
    kernel_s = ev_ms / 1e3 / steps
    bpn = BYTES_PER_NODE[args.precision]
    achieved = bpn * N / kernel_s / 1e9
    out = {
        "metric": f"MLUPS ({args.precision} TRT+MDF step, Re=100 cylinder 1601x641)",
        "value": round(mlups, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": round(wall / steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 (deviation storage; FP64 IBM)",
        "data": "synthetic (seeded 1e-3 inflow perturbation, seed 1234)",
        "config": {"workload": "fixed cylinder Re=100, 1601x641, D=40, 126 markers, TRT "
                               "Lambda=1/4, MDF [5,20] tol 1e-4, Peskin-4, U=0.04",
                   "precision": args.precision, "launch": mode, "graph": graph,
                   "nodes": N, "check_interval": ci, "Cd_last": cd, "Cl_last": cl,
                   "mdf_iter_last": iters[-1] if iters else None},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None,
                     "kernel": KERNEL_OF_MODE[mode], "bytes_per_node": bpn,
                     "avg_launch_us": round(kernel_s * 1e6, 3)},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
