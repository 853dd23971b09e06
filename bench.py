"""Benchmark of the MI355X SolverFT step.

Workload (one GPU): Cahn-Hilliard / Model B 2D, 4096x4096 periodic, h = 1,
dt = 0.01, Stencil2d2h<5,6,13>, FP32 fields, FP64 mass/energy sums, initial
psi uniform in [-0.1, 0.1) from a seed.  One bench "step" is one forward-Euler
step of the whole field.  The field is resident in HBM before timing starts.

Prints one JSON line: throughput in Gsite-updates/s, the HBM roofline of the
dominant kernel (algorithmic 8 B per site-update: psi read once, psi' written
once), accuracy against an FP64 GPU run of the same steps, FP32 mass drift,
and the CPU baseline (the FP64 C restatement of SolverFT on this host's cores).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor figure
METRIC = "Gsite-updates/s & HBM GB/s vs roofline; surrogate speedup at 1e-3 rel L2"
N = 4096
SEED = 20261202


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="target length of each CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--steps-per-launch", type=int, default=0)
    ap.add_argument("--rows-per-tile", type=int, default=0)
    return ap.parse_args()


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(args):
    """Time the FP64 oracle (OpenMP, then 1 thread) on a bounded number of
    steps of the same 4096^2 Model B workload."""
    import numpy as np

    from oracle import pyoracle
    from symphas_amd import abi, inputs, models

    desc = models.model_b((N, N), dtype=abi.F64, path=abi.PATH_REFERENCE)
    psi = inputs.uniform((N, N), seed=SEED)
    g = pyoracle.pad(desc, psi)
    lib = pyoracle.load()
    dframe = np.zeros_like(g)
    import ctypes
    dptr = ctypes.cast(ctypes.byref(desc), ctypes.c_void_p)
    gp, fp = g.ctypes.data_as(ctypes.c_void_p), dframe.ctypes.data_as(ctypes.c_void_p)
    lib.orc_refresh(dptr, gp)

    def one_step(nt):
        lib.orc_equation(dptr, gp, fp, nt)
        lib.orc_euler(dptr, gp, fp, nt)
        lib.orc_refresh(dptr, gp)

    out = {}
    for label, nt in (("omp", host_cores()), ("single", 1)):
        t0 = time.perf_counter()
        one_step(nt)
        t1 = time.perf_counter() - t0
        budget = args.cpu_seconds if label == "omp" else args.cpu_seconds / 2
        nsteps = max(1, int(budget / max(t1, 1e-6)))
        t0 = time.perf_counter()
        for _ in range(nsteps):
            one_step(nt)
        dt = time.perf_counter() - t0
        out[label] = (N * N * nsteps / dt / 1e9, nt, nsteps)
        print(f"[cpu] {label}: {nt} threads, {nsteps} steps, {out[label][0]:.4f} Gsite/s",
              file=sys.stderr, flush=True)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}), flush=True)
        sys.exit(2)

    import numpy as np
    import torch

    from symphas_amd import abi, inputs, models
    from symphas_amd.solver import FDPlan

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}), flush=True)
        sys.exit(1)
    torch.cuda.set_device(0)
    stream = torch.cuda.Stream()

    psi0 = inputs.uniform((N, N), seed=SEED, dtype=np.float32)
    d32 = models.model_b((N, N), h=1.0, dt=0.01, dtype=abi.F32, path=abi.PATH_FUSED)
    plan = FDPlan(d32)
    if args.steps_per_launch or args.rows_per_tile:
        plan.tune(args.steps_per_launch, args.rows_per_tile, 0)
    plan.set_stream(stream.cuda_stream)
    plan.upload(psi0)
    m0, e0 = plan.reduce()
    path, spl = plan.query()

    # warmup (also captures the step graphs)
    plan.step(args.warmup)
    plan.sync()
    plan.upload(psi0)
    torch.cuda.synchronize()

    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(stream):
        start.record(stream)
        plan.step(args.steps)
        stop.record(stream)
    stream.synchronize()
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop)

    sites = N * N
    value = sites * args.steps / (ms / 1e3) / 1e9
    launches = args.steps // spl + (1 if args.steps % spl else 0)
    avg_launch_s = ms / 1e3 / launches
    alg_bytes_per_launch = 8.0 * sites * spl
    achieved = alg_bytes_per_launch / avg_launch_s / 1e9

    m1, e1 = plan.reduce()
    out32 = plan.download()
    plan.close()

    accuracy = None
    if not args.no_accuracy:
        d64 = models.model_b((N, N), h=1.0, dt=0.01, dtype=abi.F64, path=abi.PATH_FUSED)
        with FDPlan(d64) as p64:
            p64.upload(psi0.astype(np.float64))
            p64.step(args.steps)
            out64 = p64.download()
        diff = out32.astype(np.float64) - out64
        accuracy = {
            "reference": "FP64 GPU run of the same steps from the same FP32-rounded field",
            "steps": args.steps,
            "max_abs_err": float(np.abs(diff).max()),
            "rel_l2_err": float(np.linalg.norm(diff) / np.linalg.norm(out64)),
        }
    drift = abs(m1 - m0) / float(np.abs(psi0.astype(np.float64)).sum())

    cpu = None
    if not args.no_cpu:
        c = cpu_baseline(args)
        cpu = {
            "value": round(c["omp"][0], 6),
            "unit": "Gsite-updates/s",
            "cores": c["omp"][1],
            "kind": "port",
            "sample": (f"FP64 C restatement of SolverFT (oracle/sps_oracle.c), OpenMP, "
                       f"{c['omp'][2]} steps of Model B 4096^2 <5,6,13> on the padded layout; "
                       f"1 thread: {c['single'][0]:.6f} Gsite-updates/s over {c['single'][2]} steps"),
            "single_thread_value": round(c["single"][0], 6),
        }

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gsite-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: psi0 uniform in [-0.1, 0.1) from SplitMix64 seed %d" % SEED,
        "config": {"workload": "Model B 2D 4096x4096 periodic, h=1, dt=0.01, "
                               "Stencil2d2h<5,6,13>, FP32 fields, FP64 sums",
                   "sites": sites, "path": "fused" if path == abi.PATH_FUSED else "reference",
                   "steps_per_launch": spl},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None,
                     "note": ("algorithmic 8 B per site-update x sites x steps per launch / "
                              "average launch time; with several steps fused per launch the "
                              "fraction can exceed 1")},
        "cpu_baseline": cpu,
        "accuracy_vs_fp64": accuracy,
        "mass_drift_rel": drift,
        "energy": {"initial": e0, "final": e1},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
