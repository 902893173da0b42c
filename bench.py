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

import numpy as np

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

# HBM bytes per k_step launch of the FP64 fused step on this workload, from rocprofv3 --pmc
# FETCH_SIZE / WRITE_SIZE (kB per dispatch), each scaled by what the counter reads for a
# 1 GiB device copy (FETCH_SIZE reads 0.5x, WRITE_SIZE 1.0x of the bytes moved); the summary
# is profiles/k_step_fp64_trt_mdf_p4_1601x641_traffic.md.
TRAFFIC_FP64_C2 = {"nodes": 1026241, "bytes": (2 * 36977.6 + 72435.2) * 1024,
                   "source": "profiles/k_step_fp64_trt_mdf_p4_1601x641_traffic.md"}


def cpu_info():
    """Host CPU as lscpu reports it: model, sockets, physical cores, hardware threads."""
    info = {"model": platform.processor() or "unknown", "sockets": None,
            "physical_cores": None, "threads": os.cpu_count() or 1}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        kv = {}
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
        info["model"] = kv.get("Model name", info["model"])
        if kv.get("Socket(s)", "").isdigit() and kv.get("Core(s) per socket", "").isdigit():
            info["sockets"] = int(kv["Socket(s)"])
            info["physical_cores"] = info["sockets"] * int(kv["Core(s) per socket"])
    except Exception:
        pass
    return info


def cpu_threads():
    """OpenMP threads for the CPU baseline: the process's CPU share (OMP_NUM_THREADS where
    set, which on a shared one-GPU host is that GPU's share of the cores), else every core
    the process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1


def cpu_baseline(seconds=15.0, threads=None):
    """The FP64 C/OpenMP oracle on the same workload: timed steps within a bounded budget."""
    from oracle import oracle as O
    info = cpu_info()
    used = O.set_threads(threads or cpu_threads())
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
                      f"step ({dt:.1f} s, {used} OpenMP threads = the process's CPU share; "
                      f"host: {info['model']}, {info['sockets']} socket(s), "
                      f"{info['physical_cores']} physical cores, {info['threads']} hardware "
                      f"threads)",
            "host": info}


def advance_aligned(sim, n, ci, graph, on_check=None):
    """Runs exactly n steps in chunks that end on multiples of ci.  A whole chunk of ci steps
    is replayed as a graph (the diagnostic sums fire on its last step); a partial chunk is
    launched step by step.  on_check() runs after each chunk that ends on a check step."""
    done = 0
    while done < n:
        c = min(ci - sim.ttt % ci, n - done)
        sim.advance(c, diag_every=ci, graph=graph and c == ci)
        done += c
        if sim.ttt % ci == 0 and on_check is not None:
            on_check()


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
    ap.add_argument("--schedule", default=None,
                    help="pin the launch schedule instead of autotuning: mode,bulk_threads,"
                         "bulk_rows,bulk_waves,step_rows,step_order[,row_layout[,stores"
                         "[,bulk_delay_us]]] (e.g. fused,256,1,0,1,blocked,tiles,normal,0), "
                         "for profiling one schedule's kernels")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default launch schedule instead of the autotuner's")
    ap.add_argument("--surrogate-steps", type=int, default=2000,
                    help="steps of the plain-vs-learned-predictor MDF comparison (0: skip)")
    ap.add_argument("--variant-steps", type=int, default=2000,
                    help="steps of the FP32-deviation comparison of an fp64 run (0: skip)")
    ap.add_argument("--shedding-steps", type=int, default=200_000,
                    help="steps of each form in --workload shedding")
    ap.add_argument("--mdf-iterations", type=int, default=0,
                    help="cylinder workload: fix the MDF sub-steps per step to this many "
                         "(minimum = maximum; 0 keeps the workload's [5, 20]), to bound what "
                         "the sub-steps cost")
    ap.add_argument("--copy-mb", default=None,
                    help="--workload copy: also plain FP32 copies of these sizes (MB of "
                         "source, comma-separated)")
    ap.add_argument("--workload", choices=["cylinder", "matrix72", "sed1", "sed64", "shedding",
                                           "copy", "large", "mdf_large"],
                    default="cylinder",
                    help="cylinder: Re=100 1601x641 (default); matrix72: the batched 72-case "
                         "Re x method x kernel x collision matrix (about 300 M nodes); sed1: "
                         "one settling disc (257x1921, CM-MRT + MDF + Peskin-4), plain MDF "
                         "against the learned predictor; sed64: "
                         "64 two-disc settling cases (193x1921, DFC + hat + CM-MRT) stepped "
                         "as one ensemble; large: the FP64 Re=200 cylinder on 3001x2401 "
                         "(1.04 GB of populations per step, beyond the last-level cache); "
                         "mdf_large: plain MDF against the learned predictor on the matrix's "
                         "Re=200 MDF cylinder (4801x1921, 377 markers, generic window), FP32 "
                         "and FP64; "
                         "copy: "
                         "a 1 GiB FP64 device copy, five times (the reference of the HBM "
                         "traffic counters and of a plain copy's bandwidth)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: the lattice fits one MI355X and this build has no "
              "multi-GPU path; ensembles would shard by instance and a single lattice by "
              "x-slabs with a three-population halo per face", file=sys.stderr)
        sys.exit(2)

    if args.workload == "matrix72":
        print(json.dumps(matrix_bench(args)), flush=True)
        return
    if args.workload == "copy":
        sizes = [float(v) for v in args.copy_mb.split(",")] if args.copy_mb else None
        print(json.dumps(copy_bench(sizes_mb=sizes)), flush=True)
        return
    if args.workload == "sed64":
        print(json.dumps(sed_bench(args)), flush=True)
        return
    if args.workload == "sed1":
        print(json.dumps(sed1_bench(args)), flush=True)
        return
    if args.workload == "shedding":
        print(json.dumps(shedding_bench(args)), flush=True)
        return
    if args.workload == "mdf_large":
        legs = {p: large_window_surrogate_leg(p, not args.no_graph) for p in ("fp32dev", "fp64")}
        r = legs["fp32dev"]
        print(json.dumps({
            "metric": "MLUPS (fp32dev TRT+MDF step, Re=200 cylinder 4801x1921, learned MDF "
                      "predictor on the generic window)",
            "value": r["MLUPS_surrogate"], "unit": "MLUPS", "n_gpus": 1,
            "steps": r["window"]["steps"], "warmup": r["window"]["start"],
            "ms_per_step": round(r["surrogate_us"] / 1e3, 5), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None,
            "dtype": "f32 (deviation storage; FP64 IBM)",
            "data": "synthetic (seeded 1e-3 inflow perturbation, seed 1234)",
            "config": {"workload": r["workload"]}, "legs": legs}), flush=True)
        return
    if args.workload == "large":
        r = large_lattice_leg(steps=max(1, args.steps // 10), warm=max(0, args.warmup // 2),
                              schedule=args.schedule)
        print(json.dumps({
            "metric": "MLUPS (fp64 DF+BGK+P4 step, Re=200 cylinder 3001x2401)",
            "value": r["MLUPS"], "unit": "MLUPS", "n_gpus": 1, "steps": r["steps"],
            "warmup": r["warmup"], "ms_per_step": round(r["device_us_per_step"] / 1e3, 5),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (uniform inflow)", "config": {"workload": r["workload"]},
            "roofline": {"bound": "hbm", "achieved": r["achieved_GBs"], "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": r["hbm_frac"], "traffic": None,
                         "kernel": KERNEL_OF_MODE[r["launch"]], "bytes_per_node": 144},
            "leg": r}), flush=True)
        return
    from iblbm_amd.solver import FixedCylinder
    cfg = workload_cfg()
    if args.mdf_iterations > 0:
        cfg.mdf_iterations = cfg.mdf_min_iterations = args.mdf_iterations
    ci = cfg.check_interval
    steps, warm = max(1, args.steps), max(0, args.warmup)
    sim = FixedCylinder(cfg, perturb=PERTURB, precision=args.precision)
    if args.serial:
        sim.ctx.set_concurrent(False)
    graph = not args.no_graph
    mode = sim.ctx.step_mode()
    N = sim.st.nodenums
    # set-up, before the warm-up: the launch schedule, then one check interval replayed as a
    # graph so that the chunk's graph is captured outside the timed region
    tuned = None
    if args.schedule:
        sim.ctx.set_schedule(**parse_schedule(args.schedule))
        tuned = sim.ctx.schedule()
    elif not args.no_tune and not args.serial:  # launch schedule from the autotuner
        tuned = sim.ctx.autotune(60)
    mode = sim.ctx.step_mode()
    advance_aligned(sim, ci, ci, graph)
    setup_steps = sim.ttt
    advance_aligned(sim, warm, ci, graph)  # exactly `warm` untimed steps
    sim.ctx.diag_fetch(0)  # allocates the pinned slots before the timed region
    sim.ctx.diag_fetch(1)
    sim.ctx.sync()
    # Exactly `steps` timed steps.  Cd and Cl every check interval, as the reference's loop
    # takes them, but read one chunk late: each chunk's sums go to a pinned slot behind its
    # steps, and the host reads the previous slot only after queueing the next chunk, so the
    # device never waits for it.
    cds = []
    st8 = {"slot": 0, "pending": None}

    def on_check():
        sim.ctx.diag_fetch(st8["slot"])
        if st8["pending"] is not None:
            cds.append(sim.cd_cl(slot=st8["pending"]))
        st8["pending"], st8["slot"] = st8["slot"], st8["slot"] ^ 1

    t0 = time.perf_counter()
    sim.ctx.mark(0)
    advance_aligned(sim, steps, ci, graph, on_check)
    sim.ctx.mark(1)
    if st8["pending"] is not None:
        cds.append(sim.cd_cl(slot=st8["pending"]))
    sim.ctx.sync()
    wall = time.perf_counter() - t0
    ev_ms = sim.ctx.elapsed_ms(0, 1)  # device timeline of the timed steps
    if not cds:  # no check step inside the timed region: Cd of the last check before it
        cds.append(sim.cd_cl())
    cd, cl = cds[-1]
    iters = [sim.ctx.mdf_stats(sim.k)["iter_count"]]
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
                               "Lambda=1/4, MDF [5,20] tol 1e-4, Peskin-4, U=0.04" +
                               (f" (MDF fixed at {args.mdf_iterations} sub-steps)"
                                if args.mdf_iterations > 0 else ""),
                   "precision": args.precision, "launch": mode, "graph": graph,
                   "schedule": tuned, "setup_steps": setup_steps,
                   "nodes": N, "check_interval": ci, "Cd_last": cd, "Cl_last": cl,
                   "mdf_iter_last": iters[-1] if iters else None},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None,
                     "kernel": KERNEL_OF_MODE[mode], "bytes_per_node": bpn,
                     "avg_launch_us": round(kernel_s * 1e6, 3)},
    }
    if N == TRAFFIC_FP64_C2["nodes"] and args.precision == "fp64" and mode == "fused":
        out["roofline"]["traffic"] = TRAFFIC_FP64_C2["bytes"]
        out["roofline"]["traffic_source"] = TRAFFIC_FP64_C2["source"]
    sim.ctx.close()
    if args.surrogate_steps > 0:
        win = dict(SURROGATE_WINDOW, steps=(max(100, args.surrogate_steps) // 100) * 100)
        out["mdf_step"] = surrogate_leg(cfg, args.precision, graph, win)
        if args.precision == "fp64" and args.variant_steps > 0:
            out["mdf_step_fp32dev"] = surrogate_leg(cfg, "fp32dev", graph, win)
    if args.precision == "fp64" and args.surrogate_steps > 0:
        out["mdf_step_large_window"] = large_window_surrogate_leg("fp32dev", graph)
    if args.precision == "fp64" and args.variant_steps > 0:
        out["persistent"] = {p: persistent_leg(cfg, p) for p in ("fp64", "fp32dev")}
    if args.precision == "fp64" and args.variant_steps > 0:
        out["fp32dev"] = fp32dev_leg(cfg, args.variant_steps, warm, graph)
        out["fp64_beyond_cache"] = large_lattice_leg()
        out["matrix72"] = dict(matrix_legs(max(10, args.steps // 200)),
                               workload=MATRIX_WORKLOAD)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(out), flush=True)


def copy_bench(reps=5, sizes_mb=None):
    """A 1 GiB FP64 device-to-device copy, `reps` times: the bytes the HBM counters must
    report for it are known exactly, which calibrates them, and its rate is a plain copy's.
    With sizes_mb, also a plain FP32 copy of each size (MB of source data; read + write is
    twice that), timed with device events over 500 copies: the rate a streaming kernel of
    the same traffic can reach, inside and beyond the 256 MB last-level cache."""
    import torch
    n = 1 << 27
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    out = {"metric": "GB/s (1 GiB FP64 device copy, read + write)",
           "value": round(2 * 8 * n / dt / 1e9, 1), "unit": "GB/s", "reps": reps,
           "ms_per_copy": round(dt * 1e3, 4)}
    del a, b
    if sizes_mb:
        by_size = {}
        for mb in sizes_mb:
            m = int(mb * 1e6) // 4
            x = torch.ones(m, dtype=torch.float32, device="cuda")
            y = torch.empty_like(x)
            for _ in range(20):
                y.copy_(x)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(500):
                y.copy_(x)
            e1.record()
            torch.cuda.synchronize()
            us = e0.elapsed_time(e1) * 1e3 / 500
            by_size[str(mb)] = {"us": round(us, 3), "traffic_MB": 2 * mb,
                                "GBs": round(2 * mb * 1e3 / us, 1),
                                "frac_of_8TBs": round(2 * mb * 1e3 / us / HBM_PEAK_GBS, 4)}
            del x, y
        out["by_size"] = by_size
    return out


def matrix_legs(steps, warm=20, plain=True):
    """The 72-case matrix in one context, FP32-deviation and FP64 storage (device-formed
    seeded initial states, launch schedule from the autotuner), and the plain FP64 path case
    by case (one context per case, default schedule), whose aggregate is the baseline of the
    batched forms."""
    from iblbm_amd.ensemble import Ensemble, case_matrix
    from iblbm_amd.solver import FixedCylinder
    cases = case_matrix()
    res = {}
    for prec in ("fp32dev", "fp64"):
        ens = Ensemble(cases, precision=prec, init="device")
        ens.advance(warm, graph=True)
        tuned = ens.autotune(10)
        us = _timed(ens, steps, True, chunk=10)
        cdcl = ens.cd_cl()
        step_s = us / 1e6
        achieved = BYTES_PER_NODE[prec] * ens.nodes / step_s / 1e9
        res[prec] = {"MLUPS": round(ens.nodes / us, 2), "ms_per_step": round(us / 1e3, 4),
                     "hbm_frac": round(achieved / HBM_PEAK_GBS, 4),
                     "schedule": tuned, "cd_finite": bool(np.isfinite(cdcl).all()),
                     "nodes": ens.nodes,
                     "bytes": round(BYTES_PER_NODE[prec] * ens.nodes / 1e9, 2)}
        ens.close()
    if plain:
        seq_nodes, seq_time = 0, 0.0
        for k, cfg in enumerate(cases):  # the plain FP64 path, one case at a time
            sim = FixedCylinder(cfg, perturb=(1000 + k, 1e-3), precision="fp64", init="device")
            sim.advance(warm, graph=True)
            us = _timed(sim, max(10, steps // 2), True, chunk=10)
            seq_nodes += sim.st.nodenums
            seq_time += us / 1e6
            sim.ctx.close()
        p = seq_nodes / seq_time / 1e6
        res["fp64_plain_case_by_case"] = {"MLUPS": round(p, 2)}
        res["speedup_fp32dev_batched_vs_fp64_plain"] = round(res["fp32dev"]["MLUPS"] / p, 3)
        res["speedup_fp64_batched_vs_fp64_plain"] = round(res["fp64"]["MLUPS"] / p, 3)
    res["steps"], res["warmup"] = steps, warm
    return res


def matrix_bench(args):
    """The batched 72-case matrix as the headline (``--workload matrix72``)."""
    steps = max(10, args.steps // 40)
    res = matrix_legs(steps, plain=args.variant_steps > 0)
    b = res["fp32dev"]
    step_s = b["ms_per_step"] / 1e3
    achieved = BYTES_PER_NODE["fp32dev"] * b["nodes"] / step_s / 1e9
    out = {
        "metric": "MLUPS (fp32dev batched 72-case fixed-cylinder matrix, one launch schedule)",
        "value": b["MLUPS"], "unit": "MLUPS", "n_gpus": 1, "steps": steps,
        "warmup": res["warmup"], "ms_per_step": b["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 (deviation storage; FP64 IBM)",
        "data": "synthetic (device-formed equilibria, hash-seeded 1e-3 velocity perturbation "
                "per case, seeds 1000..1071)",
        "config": {"workload": MATRIX_WORKLOAD, "nodes": b["nodes"], "schedule": b["schedule"],
                   "cd_finite": b["cd_finite"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_ibm || k_bulk, bytes over the whole step time",
                     "bytes_per_node": BYTES_PER_NODE["fp32dev"]},
        "matrix72": res,
    }
    return out


MATRIX_WORKLOAD = ("72 fixed cylinders: Re {20,40,100,200} x {DF,MDF,DFC} x {hat,P4} x "
                   "{BGK,TRT,CM-MRT}; Ny 641/961/1281/1921 by Re, Nx = 2.5 (Ny - 1) + 1, "
                   "D = (Ny - 1) / 16, U = 0.04")


def parse_schedule(text):
    """--schedule text -> Context.set_schedule keywords."""
    f = text.split(",")
    order = f[5] if f[5] in ("stride", "blocked", "queue") else int(f[5])
    return dict(mode=f[0], bulk_threads=int(f[1]), bulk_rows=int(f[2]), bulk_waves=int(f[3]),
                step_rows=int(f[4]), step_order=order,
                row_layout=f[6] if len(f) > 6 else "tiles",
                stores=f[7] if len(f) > 7 else "normal",
                bulk_delay_us=float(f[8]) if len(f) > 8 else 0.0)


def large_lattice_leg(steps=400, warm=200, schedule=None):
    """FP64 on one lattice far beyond the 256 MB last-level cache: the reference's Re = 200
    fixed cylinder (``scenarios/steady.py``: NN = 2401, xmax = 1.25 -> 3001 x 2401, D = 60,
    DF + BGK + Peskin-4), 1.04 GB of populations per step.  Time per step on the device
    timeline (HIP events around exactly `steps` graph-replayed steps)."""
    from iblbm_amd import SimConfig
    from iblbm_amd.solver import FixedCylinder
    cfg = SimConfig(Re=200.0, NN=2401, xmax=1.25, ymax=1.0, cylinder_center=(0.5, 0.5),
                    cylinder_D_ratio=1.0 / 40.0, lattice_u=0.1, inflow_u=0.1,
                    bc_type="dirichlet_neumann", delta_type="peskin4pt", ibm_method="DF",
                    collision_model="BGK", max_steps=60_000, check_interval=200,
                    use_convergence=False, marker_spacing_factor=2.0 / 3.0,
                    nx_formula="standard")
    sim = FixedCylinder(cfg)
    if schedule:
        sim.ctx.set_schedule(**parse_schedule(schedule))
        tuned = sim.ctx.schedule()
    else:
        tuned = sim.ctx.autotune(20)
    ci = cfg.check_interval
    advance_aligned(sim, warm, ci, True)
    sim.ctx.sync()
    t0 = time.perf_counter()
    sim.ctx.mark(0)
    advance_aligned(sim, steps, ci, True)
    sim.ctx.mark(1)
    sim.ctx.sync()
    wall = time.perf_counter() - t0
    ev_s = sim.ctx.elapsed_ms(0, 1) / 1e3 / steps
    N = sim.st.nodenums
    achieved = BYTES_PER_NODE["fp64"] * N / ev_s / 1e9
    out = {"workload": f"fixed cylinder Re=200, {sim.st.nx}x{sim.st.ny}, D=60, DF + BGK + "
                       "Peskin-4, FP64", "nodes": N,
           "bytes_per_step_GB": round(BYTES_PER_NODE["fp64"] * N / 1e9, 3),
           "steps": steps, "warmup": warm, "MLUPS": round(N * steps / wall / 1e6, 2),
           "device_us_per_step": round(ev_s * 1e6, 3), "launch": sim.ctx.step_mode(),
           "schedule": tuned, "achieved_GBs": round(achieved, 1),
           "hbm_frac": round(achieved / HBM_PEAK_GBS, 4)}
    sim.ctx.close()
    return out


def shedding_bench(args):
    """The Re = 100 cylinder (1601x641, TRT + MDF + Peskin-4, seeded inflow perturbation) run
    into periodic shedding in three forms from the same initial state: FP64 plain MDF,
    FP32-deviation storage, and FP64 with the learned force predictor (trained on the device
    on plain-MDF samples of the first 20,000 steps of its own run, then in learned mode).
    Cd and Cl every 100 steps; reported per form: the mean Cd, the Cl amplitude and the
    Strouhal number over the second half of the run (the reference's definition), the time
    per step, and each form's relative difference from FP64 plain MDF."""
    from iblbm_amd.diagnostics import compute_strouhal
    from iblbm_amd.solver import FixedCylinder
    ci = 100
    n = max(20_000, (args.shedding_steps // ci) * ci)

    def st_crossings(cl, steps, D, U):
        """St from the mean period between upward zero crossings of the mean-free Cl (linear
        interpolation between samples): far finer than the FFT bin of the same record."""
        c = np.asarray(cl) - np.mean(cl)
        t = np.asarray(steps, dtype=float)
        up = [t[k] - c[k] * (t[k + 1] - t[k]) / (c[k + 1] - c[k])
              for k in range(len(c) - 1) if c[k] < 0.0 <= c[k + 1]]
        if len(up) < 3:
            return None
        return D / (U * float(np.mean(np.diff(up))))

    res = {}
    for form in ("fp64", "fp32dev", "fp64_learned"):
        prec = "fp32dev" if form == "fp32dev" else "fp64"
        sim = FixedCylinder(workload_cfg(), perturb=PERTURB, precision=prec)
        sim.ctx.autotune(10)
        cl, cd, steps = [], [], []
        t_run = 0.0

        def run(k):
            nonlocal t_run
            for _ in range(k // ci):
                t0 = time.perf_counter()
                sim.advance(ci, diag_every=ci, graph=True)
                c, l = sim.cd_cl()
                t_run += time.perf_counter() - t0
                cd.append(c)
                cl.append(l)
                steps.append(sim.ttt)

        if form == "fp64_learned":
            pred = sim.predictor
            run(10_000)
            pred.mode("collect", capacity=1 << 20)
            run(10_000)
            pred.train(steps=3000, lr=2e-3, seed=11)
            pred.mode("learned")
            pred.stats(reset=True)
            run(n - 20_000)
            sub = pred.stats()["substeps_per_step"]
        else:
            run(n)
            sub = None
        half = len(cl) // 2
        D, U = 2.0 * sim.st.lattice_r, sim.st.in_u
        st = compute_strouhal(np.array(cl), n, D, U, ci)  # second half, FFT peak
        res[form] = {"St": st, "St_crossings": st_crossings(cl[half:], steps[half:], D, U),
                     "Cd_mean": float(np.mean(cd[half:])),
                     "Cl_amplitude": float(0.5 * (np.max(cl[half:]) - np.min(cl[half:]))),
                     "us_per_step": round(t_run * 1e6 / n, 3), "substeps_learned": sub}
        sim.ctx.close()
    base = res["fp64"]
    for form in ("fp32dev", "fp64_learned"):
        r = res[form]
        r["dSt_rel"] = abs(r["St"] - base["St"]) / abs(base["St"]) if base["St"] else None
        r["dSt_crossings_rel"] = (abs(r["St_crossings"] - base["St_crossings"]) /
                                  base["St_crossings"]
                                  if r["St_crossings"] and base["St_crossings"] else None)
        r["dCd_mean_rel"] = abs(r["Cd_mean"] - base["Cd_mean"]) / abs(base["Cd_mean"])
        r["dCl_amplitude_rel"] = (abs(r["Cl_amplitude"] - base["Cl_amplitude"]) /
                                  abs(base["Cl_amplitude"]))
    return {"metric": "Strouhal number and mean Cd, FP64 plain vs FP32-deviation vs learned "
                      "MDF (Re=100 cylinder)",
            "value": base["St"], "unit": "St", "n_gpus": 1, "steps": n, "warmup": 0,
            "ms_per_step": round(base["us_per_step"] / 1e3, 5), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64 / f32 deviation",
            "data": "synthetic (seeded 1e-3 inflow perturbation, seed 1234)",
            "config": {"workload": "fixed cylinder Re=100, 1601x641, D=40, 126 markers, TRT "
                                   "Lambda=1/4, MDF [5,20] tol 1e-4, Peskin-4, U=0.04; Cd/Cl "
                                   "every 100 steps, St over the second half (FFT peak, and "
                                   "mean zero-crossing period)"},
            "forms": res}


def sed1_cfg():
    """One settling disc: D = 32 (101 markers), density ratio 1.5, channel 8D x 60D (257x1921)
    with no-slip walls, CM-MRT + MDF [5, 20] tol 1e-4 + Peskin-4, explicit-history Verlet with
    rotation; nu = 0.02 and g = 4e-5, for a terminal Reynolds number near 40."""
    from iblbm_amd import SimConfig
    return SimConfig(Re=38.4, NN=257, xmax=1.0, ymax=7.5, cylinder_center=(0.5, 7.0),
                     cylinder_D_ratio=0.125, lattice_u=0.024, rho_ratio=1.5, gravity=4e-5,
                     motion_type="sedimentation", bc_type="settling_channel",
                     collision_model="CM_MRT", ibm_method="MDF", delta_type="peskin4pt",
                     mdf_iterations=20, mdf_min_iterations=5, mdf_tolerance=1e-4,
                     marker_spacing_factor=1.0, settling_inertia_model="explicit_history",
                     enable_rotation=True, max_steps=200_000, check_interval=200)


def sed1_bench(args):
    """The settling disc in FP64: plain MDF against the learned force predictor trained on
    this run's own MDF iterates.  After 1,000 steps from rest, 1,000 plain steps record
    samples and the network trains on the device; then both forms run 1,000 steps from one
    snapshot.  The device time of each step launch (HIP events; the markers move every step,
    so each step is set up and read back by the host) and the whole step are timed."""
    from iblbm_amd.sedimentation import Sedimentation
    cfg = sed1_cfg()
    sim = Sedimentation(cfg, precision=args.precision)
    pred = sim.predictor
    n = max(200, args.surrogate_steps // 2) if args.surrogate_steps > 0 else 1000
    for _ in range(1000):
        sim.step()
    pred.mode("collect", capacity=1 << 20)
    for _ in range(1000):
        sim.step()
    losses = pred.train(steps=3000, lr=2e-3, seed=3)
    snap = sim.snapshot()
    F0 = pred.previous_force(sim.particles[0].Lb)
    res = {}
    for mode in ("off", "learned"):
        sim.restore(snap)
        pred.set_previous_force(F0)
        pred.mode(mode)
        pred.stats(reset=True)
        dev_ms = 0.0
        t0 = time.perf_counter()
        for _ in range(n):  # Sedimentation.step with the launch bracketed by events
            sim.ttt += 1
            sim._half_step()
            sim._push_markers()
            sim.ctx.mark(0)
            sim.ctx.step(1, diag_every=1 if sim.ttt % cfg.check_interval == 0 else 0)
            sim.ctx.mark(1)
            forces = sim._forces()
            dev_ms += sim.ctx.elapsed_ms(0, 1)
            sim._full_step(forces)
        wall = time.perf_counter() - t0
        p0 = sim.particles[0]
        st = pred.stats()
        ms = sim.ctx.mdf_stats(sim.k)
        # the step kernel alone: markers held where they are, 400 steps replayed back to back
        # in graphs (the per-step launches above each carry the host's set-up and read-back)
        k_snap = sim.snapshot()
        sim.ctx.step(50, graph=True)
        sim.ctx.mark(2)
        sim.ctx.step(400, graph=True)
        sim.ctx.mark(3)
        sim.ctx.sync()
        kernel_us = sim.ctx.elapsed_ms(2, 3) * 1e3 / 400
        sim.restore(k_snap)
        res[mode] = {"kernel_us": round(kernel_us, 3), "launch_us": round(dev_ms * 1e3 / n, 3),
                     "whole_step_us": round(wall * 1e6 / n, 2),
                     "vy": float(p0.vel[1]), "y": float(p0.pos[1]),
                     "substeps_per_step": round(st["substeps_per_step"], 3),
                     "restarts": st["restarts"], "mode": sim.ctx.step_mode(),
                     "last_mdf_iter": ms["iter_count"], "last_residual": ms["last_residual"]}
    pl, le = res["off"], res["learned"]
    nu = (sim.st.tau - 0.5) / 3.0
    re_t = abs(pl["vy"]) * 2.0 * sim.st.lattice_r / nu
    nodes = sim.st.nx * sim.st.ny
    return {
        "metric": f"MDF step time, plain vs learned predictor ({args.precision} settling disc)",
        "value": round(nodes / pl["kernel_us"], 2), "unit": "MLUPS", "n_gpus": 1, "steps": n,
        "warmup": 2000, "ms_per_step": round(pl["whole_step_us"] / 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 (deviation storage; FP64 IBM)",
        "data": "synthetic (disc released from rest)",
        "config": {"workload": "settling disc D=32 (101 markers), rho_ratio 1.5, channel "
                               "8D x 60D (257x1921), CM-MRT + MDF [5,20] tol 1e-4 + Peskin-4, "
                               "explicit-history Verlet with rotation, nu=0.02, g=4e-5",
                   "nodes": nodes, "launch": pl["mode"]},
        "plain": pl, "learned": le,
        "kernel_speedup": round(pl["kernel_us"] / le["kernel_us"], 4),
        "launch_speedup": round(pl["launch_us"] / le["launch_us"], 4),
        "dvy_rel": abs(le["vy"] - pl["vy"]) / abs(pl["vy"]),
        "Re_particle": round(re_t, 3),
        "train": {"samples": pred.samples(), "adam_steps": 3000,
                  "loss_first100": float(losses[:100].mean()),
                  "loss_last100": float(losses[-100:].mean())},
    }


def sed_bench(args):
    """64 two-disc settling cases (193x1921 each, discs D = 24 of density ratio 1.01, seeded
    gap and offset; DFC + hat + CM-MRT, explicit-history Verlet with rotation) in FP32-
    deviation storage.  A step is the whole job: the vectorised particle half step of all
    128 discs, the batched marker upload, the step launch, the batched force read and the
    full step; the wall clock around K such steps gives ``value``.  Two forms: one context
    for all members, and the members split over two contexts whose steps are pipelined (the
    host moves one half's particles while the device steps the other).  The step launch
    alone (one context, markers held) is timed with HIP events for the roofline.  Also
    timed: the one-context ensemble in FP64 and the plain FP64 path case by case."""
    from iblbm_amd.sed_ensemble import PipelinedEnsembles, SedimentationEnsemble, two_disc_cfgs
    from iblbm_amd.sedimentation import Sedimentation
    n_inst = 64
    ci = 200
    cfgs = two_disc_cfgs(n_inst, seed=2024, check_interval=ci)
    steps = max(ci, (args.steps // 4 // ci) * ci)
    warm = 40
    legs = [("fp32dev", 1), ("fp32dev", 2)]
    if args.variant_steps > 0:
        legs.append(("fp64", 1))
    res = {}
    for prec, parts in legs:
        if parts == 1:
            ens = SedimentationEnsemble(cfgs, precision=prec)
            ens.run(warm)
            tuned = ens.ctx.autotune(10) if not args.no_tune else ens.ctx.schedule()
        else:
            ens = PipelinedEnsembles(cfgs, parts=parts, precision=prec)
            ens.run(warm)
            tuned = ens.autotune(10) if not args.no_tune else [e.ctx.schedule()
                                                               for e in ens.parts]
        ens.sync() if parts > 1 else ens.ctx.sync()
        t0 = time.perf_counter()
        ens.run(steps)
        wall = time.perf_counter() - t0
        leg = {"MLUPS": round(ens.nodes * steps / wall / 1e6, 2),
               "ms_per_step": round(wall * 1e3 / steps, 4), "contexts": parts,
               "nodes": ens.nodes, "schedule": tuned}
        if parts == 1:
            n_k = 200  # the step launch alone, markers held where they are
            ens.ctx.mark(0)
            ens.ctx.step(n_k)
            ens.ctx.mark(1)
            ens.ctx.sync()
            leg["step_launch_us"] = round(ens.ctx.elapsed_ms(0, 1) * 1e3 / n_k, 2)
            leg["mode"] = ens.ctx.step_mode()
            re_hist = []
            for i in range(len(ens.histories[0])):
                re = []
                for mi, h in enumerate(ens.histories):
                    if i < len(h):
                        nu = (ens.members[mi].st.tau - 0.5) / 3.0
                        re += [float(np.hypot(r["vx"], r["vy"])) * 24.0 / nu for r in h[i]]
                re_hist.append([ens.histories[0][i][0]["step"], round(float(np.mean(re)), 4),
                                round(float(np.max(re)), 4)])
            leg["re_history"] = re_hist
            leg["member0_discs_x_y_vx_vy"] = [
                [round(float(v), 6) for v in ens.trajectory_row(0)[:4]],
                [round(float(v), 6) for v in ens.trajectory_row(0)[9:13]]]
            leg["stopped"] = sum(r is not None for r in ens.reasons)
            leg["vy_all"] = [float(ens.bank.vel[k, 1]) for k in range(ens.bank.P)]
        else:
            leg["whole_step_launches_us"] = round(ens.step_launch_us(100), 2)
        res[(prec, parts)] = leg
        ens.close()
    plain = None
    if args.variant_steps > 0:  # the plain FP64 path, one case at a time, 8 of the 64 cases
        cache, nodes, t_all = {}, 0, 0.0
        for cfg in cfgs[:8]:
            sim = Sedimentation(cfg, precision="fp64", f0_cache=cache)
            for _ in range(20):
                sim.step()
            sim.ctx.sync()
            t0 = time.perf_counter()
            for _ in range(100):
                sim.step()
            sim.ctx.sync()
            t_all += time.perf_counter() - t0
            nodes += sim.st.nx * sim.st.ny * 100
            sim.ctx.close()
        plain = nodes / t_all / 1e6
    one, two = res[("fp32dev", 1)], res[("fp32dev", 2)]
    best = two if two["MLUPS"] > one["MLUPS"] else one
    achieved = BYTES_PER_NODE["fp32dev"] * one["nodes"] / (one["step_launch_us"] * 1e-6) / 1e9
    out = {
        "metric": "MLUPS (fp32dev batched 64-member two-disc settling ensemble, whole step "
                  "incl. host particle update)",
        "value": best["MLUPS"], "unit": "MLUPS", "n_gpus": 1, "steps": steps, "warmup": warm,
        "ms_per_step": best["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 (deviation storage; FP64 IBM and particles)",
        "data": "synthetic (fluid at rest; seeded gaps in [2D,4D] and offsets in [0,0.1D], "
                "numpy default_rng(2024))",
        "config": {"workload": "64 x two discs D=24, rho_ratio 1.01, channel 8D x 80D "
                               "(193x1921), DFC + hat + CM-MRT, explicit-history Verlet "
                               "with rotation, g=1.5e-3, nu=0.04",
                   "nodes": one["nodes"], "contexts": best["contexts"], "mode": one["mode"],
                   "stopped": one["stopped"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "step launch (k_ibm || k_bulk), one context, markers held",
                     "bytes_per_node": BYTES_PER_NODE["fp32dev"],
                     "avg_launch_us": one["step_launch_us"]},
        "one_context": {k: v for k, v in one.items() if k != "re_history"},
        "two_contexts_pipelined": two,
        "re_history": one["re_history"],
    }
    if ("fp64", 1) in res:
        r64 = res[("fp64", 1)]
        out["fp64_batched"] = {k: v for k, v in r64.items() if k not in ("re_history", "vy_all")}
        v32, v64 = np.array(one["vy_all"]), np.array(r64["vy_all"])
        # the same members and steps in both precisions: settling velocity of all 128 discs
        out["fp32dev_vs_fp64_settling_velocity_rel"] = float(np.max(np.abs(v32 - v64) /
                                                                    np.abs(v64)))
    out["one_context"].pop("vy_all", None)
    if plain is not None:
        out["fp64_plain_case_by_case"] = {"MLUPS": round(plain, 2), "cases": 8, "steps": 100}
        out["speedup_vs_fp64_plain"] = round(best["MLUPS"] / plain, 3)
    return out


def _timed(sim, n, graph, chunk=50):
    """Time per step (us) of n steps replayed in chunks whose graph is captured first
    (chunk + n + 2 steps run in all); the diagnostics are taken on the last step, after the
    timed region."""
    sim.advance(chunk, graph=graph)
    sim.ctx.sync()
    reps = max(1, n // chunk)
    t0 = time.perf_counter()
    for _ in range(reps):
        sim.advance(chunk, graph=graph)
    sim.ctx.sync()
    us = (time.perf_counter() - t0) / (reps * chunk) * 1e6
    sim.advance(2, diag_every=sim.ctx.step_counter() + 2)
    return us


# The learned-predictor comparison runs over a fixed window of the flow, independent of the
# headline's --steps: training samples from steps 6000..8000 (plain MDF), then both forms from
# the step-8000 state for 2000 steps, Cd sampled every 100 steps.  By step 6000 the start-up
# pressure waves have crossed the channel several times (its length is 1601 nodes, the sound
# speed 1/sqrt(3)), so the window is past the acoustic transient.
SURROGATE_WINDOW = {"train_from": 6000, "start": 8000, "steps": 2000, "sample_every": 100}
DCD_GATE = 1e-3  # |dCd| <= 0.1 % of plain MDF's Cd at every sample of the window


def surrogate_leg(cfg, precision, graph, window=SURROGATE_WINDOW):
    """MDF step time, plain against the learned force predictor, from one state of a fresh
    run of the workload in the given storage precision.

    The predictor is trained on the device on the plain-MDF steps of the window's training
    span; both forms then run the window from the same populations and previous forces.
    Reported: time per step (graph replay, no host reads), true MDF sub-steps per step, the
    largest and the mean relative Cd difference over the window's samples, each form's last
    marker slip residual, and whether the gates hold (every sample within 0.1 %, the learned
    form's slip within plain MDF's tolerance)."""
    from iblbm_amd.solver import FixedCylinder
    sim = FixedCylinder(cfg, perturb=PERTURB, precision=precision)
    tuned = sim.ctx.autotune(20)
    pred = sim.predictor
    n, every = window["steps"], window["sample_every"]
    sim.advance(window["train_from"], graph=graph)
    pred.mode("collect", capacity=1 << 19)
    sim.advance(window["start"] - window["train_from"], graph=graph)
    losses = pred.train(steps=3000, lr=2e-3, seed=11)
    f0 = sim.ctx.download_f(sim.k)
    F0 = pred.previous_force(sim.st.Lb)
    res = {}
    for mode in ("off", "learned"):
        # timing pass
        sim.ctx.upload_f(sim.k, f0)
        pred.mode(mode)
        if mode != "off":
            pred.set_previous_force(F0)
        us = _timed(sim, n, graph)
        # sampling pass over the same window from the same state
        sim.ctx.upload_f(sim.k, f0)
        pred.mode(mode)
        if mode != "off":
            pred.set_previous_force(F0)
            pred.stats(reset=True)
        cds = []
        for _ in range(n // every):
            sim.advance(every, diag_every=sim.ctx.step_counter() + every, graph=graph)
            cds.append(sim.cd_cl()[0])
        st = pred.stats() if mode != "off" else None
        res[mode] = {"us": us, "cd": np.array(cds),
                     "sub": (st["substeps_per_step"] if st
                             else float(sim.ctx.mdf_stats(sim.k)["iter_count"])),
                     "restarts": st["restarts"] if st else 0,
                     "slip": sim.ctx.mdf_stats(sim.k)["last_residual"]}
    pred.mode("off")
    sim.ctx.close()
    p, l = res["off"], res["learned"]
    rel = np.abs(l["cd"] - p["cd"]) / np.abs(p["cd"])
    tol = float(cfg.mdf_tolerance)
    return {"precision": precision, "window": dict(window), "schedule": tuned,
            "plain_us": round(p["us"], 3), "surrogate_us": round(l["us"], 3),
            "speedup": round(p["us"] / l["us"], 4), "substeps_plain_last": p["sub"],
            "substeps_surrogate": round(l["sub"], 3), "restarts": l["restarts"],
            "dCd_rel_max": float(rel.max()), "dCd_rel_mean": float(rel.mean()),
            "Cd_plain_last": float(p["cd"][-1]), "Cd_surrogate_last": float(l["cd"][-1]),
            "slip_residual_plain": p["slip"], "slip_residual_surrogate": l["slip"],
            "gates": {"dCd_rel_max_le_1e-3": bool(rel.max() <= DCD_GATE),
                      "slip_within_mdf_tolerance": bool(l["slip"] <= tol),
                      "faster_than_plain": bool(l["us"] < p["us"])},
            "train": {"samples": int(window["start"] - window["train_from"]),
                      "adam_steps": len(losses),
                      "loss_first100": float(losses[:100].mean()),
                      "loss_last100": float(losses[-100:].mean())}}


# The largest MDF instance of the 72-case matrix run alone: its 377-marker window exceeds the
# fused path and takes the generic window kernel beside the bulk kernel, and in FP32 storage
# the window is the longer of the two.  Training from step 2000 to 3000, then both forms for
# 1000 steps from the step-3000 state.
LARGE_WINDOW = {"train_from": 2000, "start": 3000, "steps": 1000, "sample_every": 100}


def large_window_surrogate_leg(precision, graph):
    """``surrogate_leg`` on the Re = 200 cylinder of the matrix (4801 x 1921, D = 120, MDF +
    Peskin-4 + TRT, 377 markers), with the window and bulk times of each form."""
    from iblbm_amd.ensemble import cylinder_cfg
    cfg = cylinder_cfg(200.0, 1921, "MDF", "peskin4pt", "TRT")
    r = surrogate_leg(cfg, precision, graph, LARGE_WINDOW)
    r["workload"] = ("fixed cylinder Re=200, 4801x1921, D=120, 377 markers, TRT, MDF [5,20] "
                     "tol 1e-4, Peskin-4 (the largest MDF instance of the 72-case matrix)")
    r["nodes"] = 4801 * 1921
    r["MLUPS_plain"] = round(r["nodes"] / r["plain_us"], 2)
    r["MLUPS_surrogate"] = round(r["nodes"] / r["surrogate_us"], 2)
    return r


def persistent_leg(cfg, precision, n=2000, chunk=200):
    """Time per step of the workload's steps in two launch forms from one state, with the
    autotuned fused schedule: per-step launches replayed as graphs of `chunk` steps (the
    headline's form) and persistent launches of `chunk` steps (one cooperative launch, a
    grid-wide barrier between steps).  Device-timed with HIP events on the context stream;
    the two forms compute the same populations bit for bit."""
    from iblbm_amd.solver import FixedCylinder
    sim = FixedCylinder(cfg, perturb=PERTURB, precision=precision)
    tuned = sim.ctx.autotune(20)
    sim.advance(chunk, graph=True)
    sim.advance(chunk, persistent=True)
    sim.ctx.sync()
    out = {"precision": precision, "steps": n, "chunk": chunk, "schedule": tuned}
    for form in ("graph", "persistent"):
        sim.ctx.mark(0)
        for _ in range(n // chunk):
            sim.advance(chunk, graph=form == "graph", persistent=form == "persistent")
        sim.ctx.mark(1)
        sim.ctx.sync()
        out[f"{form}_us_per_step"] = round(sim.ctx.elapsed_ms(0, 1) * 1e3 / n, 3)
    out["persistent_vs_graph"] = round(out["graph_us_per_step"] / out["persistent_us_per_step"],
                                       4)
    sim.ctx.close()
    return out


def fp32dev_leg(cfg, n, warm, graph):
    """The same workload with FP32-deviation storage from the same initial state: MLUPS and
    the Cd / Cl difference against the FP64 run after the same number of steps."""
    from iblbm_amd.solver import FixedCylinder
    sim = FixedCylinder(cfg, perturb=PERTURB, precision="fp32dev")
    sim.advance(warm, graph=graph)
    tuned = sim.ctx.autotune(60)
    us = _timed(sim, n, graph)
    cd32, cl32 = sim.cd_cl()
    t = sim.ttt
    ref_sim = FixedCylinder(cfg, perturb=PERTURB, precision="fp64")
    ref_sim.advance(t, diag_every=t, graph=graph)
    cd64, cl64 = ref_sim.cd_cl()
    ref_sim.ctx.close()
    out = {"MLUPS": round(sim.st.nodenums / us, 2), "ms_per_step": round(us / 1e3, 5),
           "launch": sim.ctx.step_mode(), "schedule": tuned, "steps": n, "compared_at_step": t,
           "dCd_rel_vs_fp64": abs(cd32 - cd64) / abs(cd64),
           "dCl_vs_fp64_over_Cd": abs(cl32 - cl64) / abs(cd64)}
    sim.ctx.close()
    return out


if __name__ == "__main__":
    main()
