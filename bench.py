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
# VALU issue limit in lane-instructions/s: the vendor FP32 vector peak of
# 157.3 TFLOP/s counts packed FMAs (2 lanes x 2 flops), so one non-packed
# VALU instruction per lane per cycle is 157.3e12 / 4 = 39.3e12 (256 CUs x 4
# SIMDs x 16 lanes x 2.4 GHz)
VALU_PEAK_LANE_OPS = 157.3e12 / 4
METRIC = "Gsite-updates/s & HBM GB/s vs roofline; surrogate speedup at 1e-3 rel L2"
N = 4096
SEED = 20261202
# step counts the configurations are stated at; errors are reported there
CFG1_STEPS, CFG2_STEPS, CFG3_STEPS, CFG4_STEPS = 1000, 10000, 2000, 5000


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
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the other configurations (3D, batched, surrogate)")
    ap.add_argument("--surrogate-iters", type=int, default=1200)
    ap.add_argument("--values-per-lane", type=int, default=0)
    ap.add_argument("--extra-only", choices=("cfg1", "cfg3", "cfg4", "cfg5", "checkpoint", "regional",
                                             "coupled"),
                    help="run only this one of the other configurations and print its result")
    ap.add_argument("--shape-sweep", action="store_true",
                    help="only run every launch shape the 2D autotuner tries (steps per "
                         "launch 1..4 x [4 values per lane x rows 8..64, 2 values per lane "
                         "x rows 16..128, each plus the CU-balanced heights], 6 launches "
                         "each, in that order) on the 4096^2 Model B field, for per-shape "
                         "profiles")
    ap.add_argument("--peak-probe", action="store_true",
                    help="measure this GPU's copy bandwidth and BF16 GEMM rate, then exit")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 256 MiB device to device 5 times (a known-size "
                         "transfer for rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--conv-layers", action="store_true",
                    help="only time every convolution layer shape of the surrogate U-Net at "
                         "batch 64 (256^2 input), 10 launches each, and print one JSON line per "
                         "layer with its TFLOP/s")
    ap.add_argument("--calib-probe", action="store_true",
                    help="only run a read-only kernel (sum of 1 GiB) and a write-only kernel "
                         "(fill of 1 GiB), 5 times each: the known-size transfers that "
                         "calibrate rocprofv3 FETCH_SIZE and WRITE_SIZE, too large to stay "
                         "in the 256 MB Infinity Cache")
    return ap.parse_args()


def conv_layers():
    """Per-layer rates of the surrogate's convolutions (BF16 MFMA, FP32
    accumulation): each U-Net layer shape at batch 64, 10 launches."""
    import torch
    from symphas_amd import nn
    torch.cuda.set_device(0)
    n = 64
    # (name, input side, in channels, out channels, epilogue)
    shapes = [("enc0 32->32", 256, 32, 32, 0), ("enc1a 32->64", 128, 32, 64, 0),
              ("enc1b 64->64", 128, 64, 64, 0), ("enc2a 64->128", 64, 64, 128, 0),
              ("enc2b 128->128", 64, 128, 128, 0), ("enc3a 128->256", 32, 128, 256, 0),
              ("enc3b 256->256", 32, 256, 256, 0), ("up3 256->128 +skip", 32, 256, 128, 1),
              ("up2 128->64 +skip", 64, 128, 64, 1), ("up1 64->32 +skip", 128, 64, 32, 1)]
    g = torch.Generator(device="cuda").manual_seed(0)
    for name, side, c, k, epi in shapes:
        x = torch.randn(n, side, side, c, device="cuda", generator=g).to(torch.bfloat16)
        w = (torch.randn(9, k, c, device="cuda", generator=g) * 0.05).to(torch.bfloat16)
        b = torch.zeros(k, device="cuda")
        skip = (torch.randn(n, 2 * side, 2 * side, k, device="cuda", generator=g)
                .to(torch.bfloat16) if epi else None)
        for _ in range(2):
            nn.conv3x3_raw(x, w, b, nn.ACT_SILU, epi, skip)
        torch.cuda.synchronize()
        a, e = _events(torch)
        a.record()
        for _ in range(10):
            nn.conv3x3_raw(x, w, b, nn.ACT_SILU, epi, skip)
        e.record()
        torch.cuda.synchronize()
        us = a.elapsed_time(e) / 10 * 1e3
        flops = 2 * 9 * c * k * side * side * n
        print(json.dumps({"layer": name, "batch": n, "side": side, "us": round(us, 1),
                          "TFLOPs": round(flops / us / 1e6, 1)}), flush=True)


def calib_probe():
    import torch
    torch.cuda.set_device(0)
    n = 256 * 1024 * 1024  # 1 GiB of FP32
    x = torch.ones(n, device="cuda")
    y = torch.empty(n, device="cuda")
    acc = torch.zeros((), device="cuda")
    for _ in range(5):
        acc += x.sum()        # reads 1 GiB, writes a few bytes
    for _ in range(5):
        y.fill_(2.0)          # writes 1 GiB, reads nothing
    torch.cuda.synchronize()
    print(json.dumps({"calib_probe_bytes": 4 * n, "repeats": 5,
                      "read_kernel": "torch sum (reduce)", "write_kernel": "torch fill_",
                      "check": float(acc.item())}), flush=True)


def copy_probe():
    import torch
    torch.cuda.set_device(0)
    n = 64 * 1024 * 1024  # 256 MiB of FP32
    x = torch.ones(n, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": 4 * n, "repeats": 5}), flush=True)


def peak_probe():
    """Measured peaks of this GPU: a 1 GiB device copy (HBM read + write) and a
    BF16 GEMM through the vendor BLAS (torch.matmul -> hipBLASLt, dense MFMA),
    each the best of several timed repeats."""
    import torch
    torch.cuda.set_device(0)
    out = {}
    n = 128 * 1024 * 1024  # 512 MiB of FP32 each way
    x = torch.ones(n, device="cuda")
    y = torch.empty_like(x)
    a, b = _events(torch)
    best = 1e9
    for _ in range(10):
        a.record(); y.copy_(x); b.record(); torch.cuda.synchronize()
        best = min(best, a.elapsed_time(b))
    out["copy_GBs"] = 2 * 4 * n / (best / 1e3) / 1e9
    del x, y
    m = 8192
    p = torch.randn(m, m, device="cuda", dtype=torch.bfloat16)
    q = torch.randn(m, m, device="cuda", dtype=torch.bfloat16)
    for _ in range(3):
        torch.matmul(p, q)
    torch.cuda.synchronize()
    best = 1e9
    for _ in range(10):
        a.record(); torch.matmul(p, q); b.record(); torch.cuda.synchronize()
        best = min(best, a.elapsed_time(b))
    out["bf16_gemm_TFLOPs"] = 2 * m ** 3 / (best / 1e3) / 1e12
    out["note"] = ("1 GiB device copy (read + write counted); 8192^3 BF16 GEMM via torch.matmul "
                   "(hipBLASLt); vendor peaks: HBM 8 TB/s, dense BF16 2.5 PFLOP/s")
    print(json.dumps({"peak_probe": out}), flush=True)


def shape_sweep():
    import numpy as np
    import torch
    from symphas_amd import abi, inputs, models
    from symphas_amd.solver import FDPlan
    torch.cuda.set_device(0)
    psi0 = inputs.uniform((N, N), seed=SEED, dtype=np.float32)
    d32 = models.model_b((N, N), h=1.0, dt=0.01, dtype=abi.F32, path=abi.PATH_FUSED)
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    order = []
    with FDPlan(d32) as plan:
        plan.upload(psi0)
        for k in (1, 2, 3, 4):
            for vec, rows_set in ((4, [8, 16, 32, 64]), (2, [16, 32, 64, 128])):
                # the autotuner's extra heights: m 4-wave workgroups per CU
                hl = (k * 2 + vec - 1) // vec
                nstrips = -(-N // ((64 - 2 * hl) * vec))
                for m in range(1, 7):
                    nch = m * cus * 4 // nstrips
                    ty = -(-N // nch) if nch >= 1 else 0
                    if 4 <= ty <= 512 and ty not in rows_set:
                        rows_set.append(ty)
                for rows in rows_set:
                    plan.tune(k, rows, 0, vec)
                    plan.step(6 * k)  # six launches, no graph
                    plan.sync()
                    order.append([k, rows, vec])
    print(json.dumps({"shape_sweep": order, "launches_per_shape": 6}), flush=True)


def profiled_counters(spl, rows, vec):
    """Counters of the same kernel and launch shape from the committed profile
    summary: HBM-side bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE,
    scaled by the read-only / write-only calibration kernels) and VALU lane
    instructions per site-update (SQ_INSTS_VALU x 64 / site-updates); {} for
    a shape it does not list."""
    path = ROOT / "profiles" / "fused2d_modelb_4096_f32.md"
    try:
        for line in path.read_text().splitlines():
            if line.startswith("machine-readable:"):
                kv = dict(x.strip().split(" = ") for x in line.split(":", 1)[1].split(";"))
                if (int(kv["steps_per_launch"]) == spl and int(kv["rows_per_tile"]) == rows
                        and int(kv["values_per_lane"]) == vec):
                    return {"traffic": float(kv["traffic_bytes_per_launch"]),
                            "valu_per_site_update": (float(kv["valu_per_site_update"])
                                                     if "valu_per_site_update" in kv else None),
                            "avg_launch_us": float(kv["avg_launch_us"])}
    except (OSError, ValueError, KeyError):
        return {}
    return {}


def _events(torch):
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def extra_cfg1(torch, stream):
    """Allen-Cahn 64x64, h=0.5, dt=0.01, <5,6,13>, FP64, 1000 steps on the
    reference-order kernel (bit-identical to the CPU restatement)."""
    import numpy as np
    from symphas_amd import abi, inputs, models
    from symphas_amd.solver import FDPlan
    d = models.allen_cahn((64, 64), h=0.5, dt=0.01, dtype=abi.F64, path=abi.PATH_REFERENCE)
    psi0 = inputs.uniform((64, 64), seed=1)
    with FDPlan(d) as p:
        p.set_stream(stream.cuda_stream)
        p.upload(psi0)
        p.step(100)
        p.sync()
        a, b = _events(torch)
        a.record(stream); p.step(1000); b.record(stream); stream.synchronize()
        ms = a.elapsed_time(b)
        p.upload(psi0)
        p.step(1000)
        ref = p.download()
    # the fused FP64 kernel against the reference-order one (which the tests
    # hold bit-identical to the CPU SolverFT restatement) after 1000 steps
    df = models.allen_cahn((64, 64), h=0.5, dt=0.01, dtype=abi.F64, path=abi.PATH_FUSED)
    with FDPlan(df) as p:
        p.set_stream(stream.cuda_stream)
        p.upload(psi0)
        p.step(1000)
        fused = p.download()
        p.upload(psi0)
        p.autotune(use_cache=False)
        p.step(1000)  # warm: captures the replayed launch graph
        p.sync()
        a, b = _events(torch)
        a.record(stream); p.step(1000); b.record(stream); stream.synchronize()
        ms_fused = a.elapsed_time(b)
        spl = p.query()[1]
    out = {"workload": "Allen-Cahn 64^2 FP64 reference-order kernel, 1000 steps",
           "ms_total": ms, "Gsite_updates_per_s": 64 * 64 * 1000 / (ms / 1e3) / 1e9,
           "fused_fp64": {"ms_total": ms_fused, "steps_per_launch": spl,
                          "Gsite_updates_per_s": 64 * 64 * 1000 / (ms_fused / 1e3) / 1e9},
           "error_vs_fp64_reference_order": _errors(fused, ref),
           "note": "launch-bound (4096 sites per launch); graph-replayed; the reference-order "
                   "kernel equals the CPU oracle bit for bit (tests/test_gpu_parity.py)"}
    try:
        # CPU baseline: the oracle's full 1000 steps on the host, which is
        # also the reference both GPU paths are checked against here
        from oracle import pyoracle as orc
        t0 = time.perf_counter()
        want = orc.unpad(d, orc.run(d, psi0, CFG1_STEPS, 1))
        t_cpu = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": 64 * 64 * CFG1_STEPS / t_cpu / 1e9,
                               "unit": "Gsite-updates/s", "cores": 1, "kind": "port",
                               "sample": "FP64 C restatement of SolverFT (oracle/sps_oracle.c), "
                                         "the full 1000 steps, 1 thread"}
        out["error_vs_cpu_oracle"] = {"steps": CFG1_STEPS,
                                      "reference_order": _errors(ref, want),
                                      "fused_fp64": _errors(fused, want)}
    except Exception as exc:  # reported, never hidden
        out["cpu_baseline"] = {"error": f"{type(exc).__name__}: {exc}"}
    return out


def _errors(got, want):
    import numpy as np
    g = np.asarray(got, dtype=np.float64)
    w = np.asarray(want, dtype=np.float64)
    return {"max_abs_err": float(np.max(np.abs(g - w))),
            "rel_l2_err": float(np.linalg.norm((g - w).ravel()) / np.linalg.norm(w.ravel()))}


def extra_cfg3(torch, stream, steps=CFG3_STEPS, cpu_seconds=6.0):
    """Model B 3D 512^3, h=1, dt=0.005, <7,10,25>, FP32 fused kernel, timed
    over the stated 2000 steps and compared there with an FP64 run; the FP64
    oracle timed on a few steps of the same field beside it (cpu_seconds = 0
    skips that)."""
    import numpy as np
    from symphas_amd import abi, inputs, models
    from symphas_amd.solver import FDPlan
    n = 512
    d = models.model_b((n, n, n), h=1.0, dt=0.005, stencil=models.Stencil3d2h(7, 10, 25),
                       dtype=abi.F32, path=abi.PATH_FUSED)
    psi = inputs.uniform((n, n, n), seed=3, lo=-0.05, hi=0.05, dtype=np.float32)
    with FDPlan(d) as p:
        p.set_stream(stream.cuda_stream)
        p.upload(psi)
        # a device copy of the initial field, made before autotuning and
        # warmup: a host-to-device transfer of the field right before the
        # timed steps slows the kernels of the next ~25 ms
        psi_dev = torch.from_numpy(psi).to("cuda")
        torch.cuda.synchronize()
        shape = p.autotune(use_cache=False)
        m0, _ = p.reduce()
        p.step(64)  # warm: the same call captures the replayed launch graph
        p.copy_from_device(psi_dev.data_ptr())  # back to the initial field
        p.sync()
        del psi_dev
        torch.cuda.synchronize()
        a, b = _events(torch)
        a.record(stream); p.step(steps); b.record(stream); stream.synchronize()
        ms = a.elapsed_time(b)
        m1, _ = p.reduce()
        got = p.download()
    # the same steps in FP64 from the same FP32-rounded field
    d64 = models.model_b((n, n, n), h=1.0, dt=0.005, stencil=models.Stencil3d2h(7, 10, 25),
                         dtype=abi.F64, path=abi.PATH_FUSED)
    with FDPlan(d64) as p:
        p.set_stream(stream.cuda_stream)
        p.upload(psi.astype(np.float64))
        p.autotune(use_cache=False)
        p.step(steps)
        want = p.download()
    err = _errors(got, want)
    del got, want
    cpu = None
    if cpu_seconds > 0:
        dcpu = models.model_b((n, n, n), h=1.0, dt=0.005, stencil=models.Stencil3d2h(7, 10, 25),
                              dtype=abi.F64, path=abi.PATH_REFERENCE)
        c = cpu_rate(dcpu, psi.astype(np.float64), (cpu_seconds, cpu_seconds / 2), "512^3")
        cpu = {"value": round(c["omp"][0], 6), "unit": "Gsite-updates/s", "cores": c["omp"][1],
               "kind": "port",
               "sample": (f"FP64 C restatement of SolverFT (oracle/sps_oracle.c), OpenMP, "
                          f"{c['omp'][2]} steps of this 512^3 field on the padded layout; "
                          f"1 thread: {c['single'][0]:.6f} Gsite-updates/s over "
                          f"{c['single'][2]} steps"),
               "single_thread_value": round(c["single"][0], 6)}
    g = n ** 3 * steps / (ms / 1e3) / 1e9
    return {"workload": "Model B 3D 512^3 FP32, Stencil3d2h<7,10,25>, dt=0.005",
            "steps": steps, "Gsite_updates_per_s": g,
            "launch_shape": {"steps_per_launch": shape[0], "rows_per_wave": shape[1],
                             "planes_per_chunk": shape[2]},
            "roofline": {"achieved_GBs": 8 * g, "peak_GBs": HBM_PEAK_GBS, "frac": 8 * g / HBM_PEAK_GBS},
            "mass_drift_rel": abs(m1 - m0) / float(np.abs(psi.astype(np.float64)).sum()),
            "cpu_baseline": cpu,
            "accuracy_vs_fp64": {"reference": "FP64 GPU run of the same steps from the same "
                                              "FP32-rounded field", "steps": steps, **err}}


def extra_cfg4(torch, stream, steps=CFG4_STEPS):
    """256 instances, half Allen-Cahn and half Model B, sides in
    {128,...,512}, M and kappa in [0.5, 2], FP32, one persistent launch per
    step group, hipGraph-replayed, timed over the stated 5000 steps from the
    initial fields and compared there with FP64 runs of every instance;
    against the same instances stepped one after another (the reference's
    find_solution pattern)."""
    import numpy as np
    from symphas_amd import abi, inputs, models
    from symphas_amd.solver import FDBatch, FDPlan
    inst = models.batched_instances(256, seed=7)
    sites = sum(d.n[0] * d.n[1] for d, _ in inst)
    fields = [inputs.uniform((d.n[1], d.n[0]), seed=s, dtype=np.float32) for d, s in inst]
    with FDBatch([d for d, _ in inst]) as bt:
        bt.set_stream(stream.cuda_stream)
        for i, f in enumerate(fields):
            bt.upload(i, f)
        # device copies of the initial fields: restoring them device to device
        # after warmup avoids a host upload right before the timed steps
        dev0 = [torch.from_numpy(f).to("cuda") for f in fields]
        torch.cuda.synchronize()
        bshape = bt.autotune()
        bt.step(48)  # warm: captures the launch graphs
        for i, t in enumerate(dev0):
            bt.copy_from_device(i, t.data_ptr())
        bt.sync()
        del dev0
        a, b = _events(torch)
        a.record(stream); bt.step(steps); b.record(stream); stream.synchronize()
        ms = a.elapsed_time(b)
        got = [bt.download(i) for i in range(len(inst))]
    g = sites * steps / (ms / 1e3) / 1e9
    # every instance in FP64 (single plans) for the same steps
    worst = {"max_abs_err": 0.0, "rel_l2_err": 0.0}
    for (d, _), f, gf in zip(inst, fields, got):
        d64 = type(d).from_buffer_copy(d)  # same descriptor, FP64 storage
        d64.dtype = abi.F64
        with FDPlan(d64) as pl:
            pl.set_stream(stream.cuda_stream)
            pl.upload(f.astype(np.float64))
            pl.step(steps)
            e = _errors(gf, pl.download())
        worst = {k: max(worst[k], e[k]) for k in worst}
    plans = []
    for (d, _), f in zip(inst, fields):
        pl = FDPlan(d)
        pl.set_stream(stream.cuda_stream)
        pl.upload(f)
        plans.append(pl)
    for pl in plans:
        pl.step(4)
    stream.synchronize()
    a, b = _events(torch)
    a.record(stream)
    for pl in plans:
        pl.step(48)
    b.record(stream); stream.synchronize()
    seq = sites * 48 / (a.elapsed_time(b) / 1e3) / 1e9
    for pl in plans:
        pl.close()
    return {"workload": "batched: 256 instances (128 Allen-Cahn, 128 Model B), sides "
                        "{128,192,256,384,512}, M, kappa ~ U[0.5,2] (Model B draws kept "
                        "within the explicit-Euler bound), FP32",
            "sites": sites, "steps": steps, "Gsite_updates_per_s": g,
            "launch_shape": {"steps_per_launch": bshape[0], "rows_per_item": bshape[1]},
            "roofline": {"achieved_GBs": 8 * g, "peak_GBs": HBM_PEAK_GBS, "frac": 8 * g / HBM_PEAK_GBS},
            "sequential_plans_Gsite_updates_per_s": seq,
            "speedup_vs_sequential": g / seq,
            "accuracy_vs_fp64": {"reference": "each instance as an FP64 single-field GPU run of "
                                              "the same steps from the same FP32 field",
                                 "steps": steps, "worst_instance": worst}}


def extra_cfg5(torch, stream, iters):
    """Surrogate: Model B 256^2 batches of 64, 50-step advances, trained on
    solver trajectories in this run; time-to-solution against the plain GPU
    solver at a 1e-3 relative L2 check with fallback."""
    import numpy as np
    from symphas_amd import surrogate as sg
    from symphas_amd.nn import UNet
    n, k, batch = 256, 50, 64
    torch.manual_seed(0)
    net = UNet(seed=1).cuda()
    # held-out fields, advanced 2000 steps by the plain solver (late times)
    psi0 = torch.stack([torch.from_numpy(sg.initial_field(n, 2 * 10**8 + i))
                        for i in range(batch)]).cuda()
    warm = sg.PlainSolver(batch, n)
    start = warm.advance(psi0, 2000)
    ref_adv = warm.advance(start, k)
    warm.close()

    def held_out(m):
        # mean single-advance relative L2 of the network on the held-out batch
        with torch.no_grad():
            return float(sg.rel_l2(m(start), ref_adv).mean())

    eval_at = sorted({max(1, iters // 8), max(1, iters // 4), max(1, iters // 2), iters})
    log = sg.train(net, iters=iters, batch=32, n=n, k=k, horizon=20000, seed=11,
                   log_every=max(1, iters // 4), verbose=False, evaluate=held_out,
                   eval_at=eval_at)
    n_adv = 20
    plain, t_plain = sg.time_plain(start, n_adv * k, n)
    ctl = sg.SurrogateSolver(net, batch, n, k=k, tol=1e-3, check_every=4)
    ctl.advance(start, 1)
    out, rep = ctl.advance(start, n_adv)
    ctl.close()
    err = sg.rel_l2(out, plain)
    t_net = sg.time_network(net, start)
    flops = net.flops_per_sample(n, n) * batch
    t_roof = net.roofline_seconds(n, n, batch)
    # surrogate alone (no checks) against the plain solver
    with torch.no_grad():
        pred = net(start)
    one = sg.PlainSolver(batch, n)
    ref1 = one.advance(start, k)
    one.close()
    e1 = sg.rel_l2(pred, ref1)
    return {"workload": "Model B 256^2, batches of 64, 20 advances of 50 steps from states "
                        "2000 steps after seeded noise + mean offset",
            "train": {"iters": iters, "batch": 32, "seconds": log.seconds,
                      "final_loss": log.losses[-1][1] if log.losses else None,
                      "held_out_single_advance_rel_l2": [
                          {"iters": i, "mean": e} for i, e in log.evaluations],
                      "note": "pairs from 32 solver trajectories generated on the fly "
                              "(random lifetimes up to 20000 steps); the held-out curve "
                              "shows where the network's accuracy levels off against the "
                              "1e-3 gate"},
            "plain_seconds": t_plain, "surrogate_with_fallback_seconds": rep.seconds,
            "speedup": t_plain / rep.seconds,
            "rel_l2_vs_plain": {"max": float(err.max()), "mean": float(err.mean())},
            "single_advance_rel_l2": {"max": float(e1.max()), "mean": float(e1.mean())},
            "checks": rep.checks, "fallback_samples": rep.fallbacks,
            "network_forward_seconds": t_net,
            "mfma": {"flops_per_forward": flops, "achieved_TFLOPs": flops / t_net / 1e12,
                     "peak_TFLOPs": 2500.0, "utilisation": flops / t_net / 2.5e15,
                     "combined_roofline_seconds": t_roof,
                     "frac_of_combined_roofline": t_roof / t_net,
                     "note": "the combined roofline sums max(FLOPs / 2.5 PFLOP/s, bytes / "
                             "8 TB/s) over the pass's kernels; the 32-channel layers at 256^2, "
                             "the upsample-add layers, the first and last layers are bound by "
                             "their HBM bytes, so utilisation cannot exceed "
                             "flops_per_forward / 2.5e15 / combined_roofline_seconds"},
            "plain_50_step_seconds_per_batch": t_plain / n_adv,
            "note": ("the specified network costs 13.4 GFLOP per sample per advance; 50 "
                     "plain steps of one 256^2 field cost ~3.3M site-updates, so the "
                     "surrogate cannot be faster on this GPU at any MFMA efficiency")}


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_rate(desc, psi, budgets, what):
    """Time the FP64 oracle's SolverFT step (equation, Euler, refresh on the
    padded layout) on a bounded number of steps: OpenMP on the host's cores,
    then 1 thread.  budgets = seconds for each; returns label -> (Gsite/s,
    threads, steps)."""
    import ctypes
    import numpy as np

    from oracle import pyoracle

    g = pyoracle.pad(desc, psi)
    lib = pyoracle.load()
    dframe = np.zeros_like(g)
    dptr = ctypes.cast(ctypes.byref(desc), ctypes.c_void_p)
    gp, fp = g.ctypes.data_as(ctypes.c_void_p), dframe.ctypes.data_as(ctypes.c_void_p)
    lib.orc_refresh(dptr, gp)
    sites = int(np.prod([desc.n[a] for a in range(desc.dim)]))

    def one_step(nt):
        lib.orc_equation(dptr, gp, fp, nt)
        lib.orc_euler(dptr, gp, fp, nt)
        lib.orc_refresh(dptr, gp)

    out = {}
    for (label, nt), budget in zip((("omp", host_cores()), ("single", 1)), budgets):
        t0 = time.perf_counter()
        one_step(nt)
        t1 = time.perf_counter() - t0
        if t1 >= budget:  # the first step already fills the budget: it is the sample
            nsteps, dt = 1, t1
        else:
            nsteps = max(1, int(budget / max(t1, 1e-6)))
            t0 = time.perf_counter()
            for _ in range(nsteps):
                one_step(nt)
            dt = time.perf_counter() - t0
        out[label] = (sites * nsteps / dt / 1e9, nt, nsteps)
        print(f"[cpu] {what} {label}: {nt} threads, {nsteps} steps, {out[label][0]:.4f} Gsite/s",
              file=sys.stderr, flush=True)
    return out


def cpu_baseline(args):
    """The FP64 oracle on a bounded number of steps of the same 4096^2 Model B
    workload."""
    from symphas_amd import abi, inputs, models

    desc = models.model_b((N, N), dtype=abi.F64, path=abi.PATH_REFERENCE)
    psi = inputs.uniform((N, N), seed=SEED)
    return cpu_rate(desc, psi, (args.cpu_seconds, args.cpu_seconds / 2), "4096^2")


def extra_checkpoint(torch, stream):
    """GNU checkpoint file (the reference's save_grid / read_block format) of
    the 4096^2 FP32 field, written and read on the GPU, against the
    reference's algorithm on the CPU (one fprintf("%.10E ") per value,
    restated in the oracle) for the same field."""
    import numpy as np
    from symphas_amd import abi, inputs, models
    from symphas_amd.solver import FDPlan
    psi = inputs.uniform((N, N), seed=SEED, dtype=np.float32)
    d32 = models.model_b((N, N), h=1.0, dt=0.01, dtype=abi.F32, path=abi.PATH_FUSED)
    with FDPlan(d32) as p, FDPlan(d32) as q:
        p.upload(psi)
        p.checkpoint_text(index=1)  # warm
        t0 = time.perf_counter()
        text = p.checkpoint_text(index=10000)
        t_write = time.perf_counter() - t0
        q.read_checkpoint(text)
        t0 = time.perf_counter()
        idx = q.read_checkpoint(text)
        t_read = time.perf_counter() - t0
        same = bool(np.array_equal(q.download(), psi)) and idx == 10000
    out = {"workload": "GNU checkpoint text of Model B 4096^2 FP32 (16.8M values, %.10E)",
           "bytes": len(text), "write_s": t_write, "read_s": t_read,
           "write_Mvalues_per_s": N * N / t_write / 1e6,
           "read_Mvalues_per_s": N * N / t_read / 1e6,
           "round_trip_exact": same,
           "note": "write = length pass + format pass + device-to-host copy into pageable memory; "
                   "read = host-to-device copy + tokenise + parse"}
    try:
        from oracle import pyoracle as orc  # CPU baseline: the reference's loop
        d64 = models.model_b((N, N), h=1.0, dt=0.01, dtype=abi.F64, path=abi.PATH_REFERENCE)
        g = orc.pad(d64, psi.astype(np.float64))
        t0 = time.perf_counter()
        ref = orc.checkpoint_text(d64, g, index=10000)
        t_cpu = time.perf_counter() - t0
        out["cpu_baseline"] = {"write_s": t_cpu, "cores": 1, "kind": "port",
                               "byte_identical": ref == text,
                               "sample": "oracle/sps_oracle.c orc_checkpoint_text, the full field"}
        out["write_speedup_vs_cpu"] = t_cpu / t_write
    except Exception as exc:  # reported, never hidden
        out["cpu_baseline"] = {"error": f"{type(exc).__name__}: {exc}"}
    out["plot_and_xdr"] = extra_plot_files(psi, d32)
    return out


def extra_plot_files(psi, d32):
    """The reference's other data files of the same field, written on the GPU:
    GNU matrix and CSV plot text ("% 10lE"), XDR (big-endian doubles; also
    read back), with the oracle's byte-at-a-time restatements timed beside
    them."""
    import numpy as np
    from symphas_amd import abi, models
    from symphas_amd.solver import FDPlan
    out = {}
    with FDPlan(d32) as p:
        p.upload(psi)
        for name, fn in (("gnu_plot", lambda: p.plot_text(abi.PLOT_GNU)),
                         ("csv", lambda: p.plot_text(abi.PLOT_CSV)),
                         ("xdr", lambda: p.checkpoint_xdr(index=10000))):
            fn()  # warm
            t0 = time.perf_counter()
            data = fn()
            t = time.perf_counter() - t0
            out[name] = {"bytes": len(data), "write_s": t,
                         "write_Mvalues_per_s": N * N / t / 1e6}
            if name == "xdr":
                p.read_checkpoint_xdr(data)
                t0 = time.perf_counter()
                p.read_checkpoint_xdr(data)
                out[name]["read_s"] = time.perf_counter() - t0
                out[name]["round_trip_exact"] = bool(np.array_equal(p.download(), psi))
            try:
                from oracle import pyoracle as orc
                d64 = models.model_b((N, N), h=1.0, dt=0.01, dtype=abi.F64,
                                     path=abi.PATH_REFERENCE)
                g = orc.pad(d64, psi.astype(np.float64))
                t0 = time.perf_counter()
                ref = (orc.xdr_write(d64, g, 10000) if name == "xdr" else
                       orc.plot_text(d64, g, abi.PLOT_GNU if name == "gnu_plot" else abi.PLOT_CSV))
                t_cpu = time.perf_counter() - t0
                out[name]["cpu_baseline"] = {"write_s": t_cpu, "cores": 1, "kind": "port",
                                             "byte_identical": ref == data}
                out[name]["write_speedup_vs_cpu"] = t_cpu / t
            except Exception as exc:  # reported, never hidden
                out[name]["cpu_baseline"] = {"error": f"{type(exc).__name__}: {exc}"}
            del data
    return out


def extra_coupled(torch, stream, steps=1000):
    """Two coupled fields (Gray-Scott and Model C, 4096^2 FP32): the fused
    two-field kernel, autotuned; 16 B of algorithmic traffic per site-update
    (both fields read once and written once)."""
    import numpy as np
    from symphas_amd import abi, inputs, models
    from symphas_amd.solver import MultiPlan
    out = {}
    u = np.ones((N, N), np.float32)
    v = np.zeros((N, N), np.float32)
    u[N // 2 - 64:N // 2 + 64, N // 2 - 64:N // 2 + 64] = 0.5
    v[N // 2 - 64:N // 2 + 64, N // 2 - 64:N // 2 + 64] = 0.25
    psi = inputs.uniform((N, N), seed=SEED, lo=-0.1, hi=0.1, dtype=np.float32)
    rho = inputs.uniform((N, N), seed=SEED + 1, lo=-0.1, hi=0.1, dtype=np.float32)
    for name, d, fields in (
            ("gray_scott_4096_fp32", models.gray_scott((N, N), dtype=abi.F32, path=abi.PATH_FUSED),
             (u, v)),
            ("model_c_4096_fp32", models.model_c((N, N), h=1.0, dt=0.01, dtype=abi.F32,
                                                 path=abi.PATH_FUSED), (psi, rho))):
        with MultiPlan(d) as p:
            p.set_stream(stream.cuda_stream)
            for f, a in enumerate(fields):
                p.upload(f, a)
            k, rows, vec = p.autotune()
            p.step(8)
            a_, b_ = _events(torch)
            a_.record(stream); p.step(steps); b_.record(stream); stream.synchronize()
            ms = a_.elapsed_time(b_)
            finite = all(bool(np.isfinite(p.download(f)).all()) for f in range(2))
        g = N * N * steps / (ms / 1e3) / 1e9
        out[name] = {"steps": steps, "Gsite_updates_per_s": g,
                     "launch_shape": {"steps_per_launch": k, "rows_per_chunk": rows,
                                      "values_per_lane": vec},
                     "roofline": {"achieved_GBs": 16 * g, "peak_GBs": HBM_PEAK_GBS,
                                  "frac": 16 * g / HBM_PEAK_GBS},
                     "finite": finite}
    # 3D: Gray-Scott and Model C on 256^3 FP32 (Stencil3d2h<7,10,25>), fused
    # two-field plane streaming against the reference-order kernel
    n3 = 256
    st3 = models.Stencil3d2h(7, 10, 25)
    u3 = np.ones((n3, n3, n3), np.float32)
    v3 = np.zeros((n3, n3, n3), np.float32)
    u3[96:160, 96:160, 96:160] = 0.5
    v3[96:160, 96:160, 96:160] = 0.25
    p3 = inputs.uniform((n3, n3, n3), seed=SEED + 2, lo=-0.1, hi=0.1, dtype=np.float32)
    r3 = inputs.uniform((n3, n3, n3), seed=SEED + 3, lo=-0.1, hi=0.1, dtype=np.float32)
    for name, mk, fields, steps3 in (
            ("gray_scott_256cubed_fp32", models.gray_scott, (u3, v3), 200),
            ("model_c_256cubed_fp32", models.model_c, (p3, r3), 100)):
        rates = {}
        for path in (abi.PATH_FUSED, abi.PATH_REFERENCE):
            kw = {"h": 1.0, "dt": 0.002} if mk is models.model_c else {}
            d = mk((n3, n3, n3), stencil=st3, dtype=abi.F32, path=path, **kw)
            with MultiPlan(d) as p:
                p.set_stream(stream.cuda_stream)
                for f, a in enumerate(fields):
                    p.upload(f, a)
                if path == abi.PATH_FUSED:
                    shape3 = p.autotune()
                nst = steps3 if path == abi.PATH_FUSED else 10
                p.step(nst)
                a_, b_ = _events(torch)
                a_.record(stream); p.step(nst); b_.record(stream); stream.synchronize()
                rates[path] = n3 ** 3 * nst / (a_.elapsed_time(b_) / 1e3) / 1e9
        g = rates[abi.PATH_FUSED]
        out[name] = {"steps": steps3, "Gsite_updates_per_s": g,
                     "launch_shape": {"steps_per_launch": shape3[0], "rows_per_wave": shape3[1]},
                     "roofline": {"achieved_GBs": 16 * g, "peak_GBs": HBM_PEAK_GBS,
                                  "frac": 16 * g / HBM_PEAK_GBS},
                     "reference_order_Gsite_updates_per_s": rates[abi.PATH_REFERENCE],
                     "speedup_vs_reference_order": g / rates[abi.PATH_REFERENCE]}
    return out


def extra_regional(torch, stream):
    """Regional grid (SolverSystemFDwSD): Model B 4096^2 FP32 whose activity
    is one disc of psi = 0.5 (radius 400 sites) on a background at 0; the box
    follows it (free resize every 0.5 time units).  Against the same
    reference-order arithmetic over the whole domain, and the fused step."""
    import numpy as np
    from symphas_amd import abi, models
    from symphas_amd.solver import FDPlan, RegionalPlan
    n, steps = 4096, 200
    yy, xx = np.mgrid[0:n, 0:n]
    psi = np.where((xx - 1500.0) ** 2 + (yy - 2300.0) ** 2 < 400.0 ** 2, 0.5, 0.0)
    psi = psi.astype(np.float32)
    params = abi.Regional(1e-6, 2.0, 0.5, 0, 0, 0.05)

    def timed(plan, k):
        a, b = _events(torch)
        a.record(stream)
        plan.step(k)
        b.record(stream)
        stream.synchronize()
        return a.elapsed_time(b) / 1e3

    d = models.model_b((n, n), h=1.0, dt=0.01, dtype=abi.F32, path=abi.PATH_REFERENCE)
    df = models.model_b((n, n), h=1.0, dt=0.01, dtype=abi.F32, path=abi.PATH_FUSED)

    def regional(desc):
        with RegionalPlan(desc, params) as rp:
            rp.upload(psi)
            rp.step(1)  # first update: cutoff, empty, first box
            rp.sync()
            t0 = time.perf_counter()  # the plan runs on its own stream
            rp.step(steps)
            rp.sync()
            info = rp.region()
            return time.perf_counter() - t0, [int(info.dims[0]), int(info.dims[1])]

    t_reg, box = regional(d)
    t_regf, _ = regional(df)
    with FDPlan(d) as p:
        p.set_stream(stream.cuda_stream)
        p.upload(psi)
        p.step(2)
        t_ref = timed(p, steps)
    with FDPlan(df) as p:
        p.set_stream(stream.cuda_stream)
        p.upload(psi)
        p.step(8)
        t_fused = timed(p, steps)
    sites = n * n
    return {"workload": "Model B 4096^2 FP32, one disc of radius 400 sites on a 0 background, "
                        "regional grid with free resize every 0.5 time units, 200 steps",
            "box_dims": box, "box_fraction_of_domain": box[0] * box[1] / sites,
            "regional_reference_order_ms_per_step": t_reg / steps * 1e3,
            "regional_fused_ms_per_step": t_regf / steps * 1e3,
            "whole_domain_reference_order_ms_per_step": t_ref / steps * 1e3,
            "whole_domain_fused_ms_per_step": t_fused / steps * 1e3,
            "speedup_reference_order": t_ref / t_reg,
            "speedup_fused": t_fused / t_regf,
            "domain_Gsite_updates_per_s_equivalent_fused": sites * steps / t_regf / 1e9,
            "note": "the regional plan steps only the box, in the reference order "
                    "(bit-identical to the CPU SolverFT in FP64) or the fused form (four steps "
                    "per launch between resizes while the box's layer is frozen); rates are "
                    "quoted per domain site for comparison with whole-domain plans (the fused "
                    "whole-domain plan also fuses 4 steps per launch)"}


def extra_configs(torch, stream, args):
    return (("cfg1_allen_cahn_64_fp64", lambda: extra_cfg1(torch, stream)),
            ("cfg3_model_b_512cubed_fp32",
             lambda: extra_cfg3(torch, stream,
                                cpu_seconds=0.0 if args.no_cpu else args.cpu_seconds / 2)),
            ("cfg4_batched_256", lambda: extra_cfg4(torch, stream)),
            ("cfg5_surrogate", lambda: extra_cfg5(torch, stream, args.surrogate_iters)),
            ("checkpoint_gnu_text", lambda: extra_checkpoint(torch, stream)),
            ("coupled_fields", lambda: extra_coupled(torch, stream)),
            ("regional_grid", lambda: extra_regional(torch, stream)))


def extra_only(args):
    """Run one of the other configurations alone (for profiling) and print its
    result as one JSON line."""
    import torch
    torch.cuda.set_device(0)
    stream = torch.cuda.Stream()
    for name, fn in extra_configs(torch, stream, args):
        if name.startswith(args.extra_only + "_"):
            with torch.cuda.stream(stream):
                print(json.dumps({name: fn()}), flush=True)
            return
    raise SystemExit(f"unknown configuration {args.extra_only}")


def main():
    args = parse()
    if args.extra_only:
        extra_only(args)
        return
    if args.peak_probe:
        peak_probe()
        return
    if args.copy_probe:
        copy_probe()
        return
    if args.calib_probe:
        calib_probe()
        return
    if args.conv_layers:
        conv_layers()
        return
    if args.shape_sweep:
        shape_sweep()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; --gpus must be 1",
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
    plan.set_stream(stream.cuda_stream)
    plan.upload(psi0)
    # the initial field also stays on the device, copied before autotuning
    # and warmup: a host-to-device transfer right before the timed steps
    # slows the kernels of the next tens of milliseconds
    psi0_dev = torch.from_numpy(psi0).to("cuda")
    torch.cuda.synchronize()
    if args.steps_per_launch or args.rows_per_tile or args.values_per_lane:
        plan.tune(args.steps_per_launch, args.rows_per_tile, 0, args.values_per_lane)
        shape = (args.steps_per_launch, args.rows_per_tile, 0, args.values_per_lane)
        tuned = "fixed by flags"
    else:
        shape = plan.autotune(use_cache=False)
        tuned = ("autotuned: %d steps per launch, %d rows per chunk, %d values per lane"
                 % (shape[0], shape[1], shape[3]))
    m0, e0 = plan.reduce()
    path, spl = plan.query()

    # warmup (also captures the step graphs)
    plan.step(args.warmup)
    plan.sync()
    # back to the initial field, device to device
    plan.copy_from_device(psi0_dev.data_ptr())
    plan.sync()
    del psi0_dev
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
    # the least HBM traffic a launch needs is one read and one write of the
    # field, whatever its steps per launch: that is the roofline's numerator
    min_bytes_per_launch = 8.0 * sites
    achieved = min_bytes_per_launch / avg_launch_s / 1e9
    unfused_equivalent = 8.0 * sites * spl / avg_launch_s / 1e9
    counters = profiled_counters(shape[0], shape[1], shape[3])
    traffic = counters.get("traffic")
    valu = counters.get("valu_per_site_update")
    m1, e1 = plan.reduce()
    out32 = plan.download()
    plan.close()

    accuracy = None
    fp64 = None
    if not args.no_accuracy:
        # errors at the workload's stated length, whatever --steps is
        if args.steps == CFG2_STEPS:
            a32 = out32
        else:
            with FDPlan(d32) as p32:
                p32.set_stream(stream.cuda_stream)
                p32.tune(shape[0], shape[1], 0, shape[3])
                p32.upload(psi0)
                p32.step(CFG2_STEPS)
                a32 = p32.download()
        d64 = models.model_b((N, N), h=1.0, dt=0.01, dtype=abi.F64, path=abi.PATH_FUSED)
        with FDPlan(d64) as p64:
            p64.set_stream(stream.cuda_stream)
            p64.upload(psi0.astype(np.float64))
            psi64_dev = torch.from_numpy(psi0.astype(np.float64)).to("cuda")
            torch.cuda.synchronize()
            s64 = p64.autotune(use_cache=False)
            p64.step(200)  # warm: captures the launch graphs
            p64.copy_from_device(psi64_dev.data_ptr())  # device to device
            p64.sync()
            del psi64_dev
            a, b = _events(torch)
            a.record(stream); p64.step(CFG2_STEPS); b.record(stream); stream.synchronize()
            ms64 = a.elapsed_time(b)
            out64 = p64.download()
        fp64 = {"Gsite_updates_per_s": sites * CFG2_STEPS / (ms64 / 1e3) / 1e9,
                "steps": CFG2_STEPS,
                "launch_shape": {"steps_per_launch": s64[0], "rows_per_tile": s64[1],
                                 "values_per_lane": s64[3]},
                "note": "the same workload in FP64 (16 B per site-update), fused kernel, "
                        "autotuned; its output is the reference of accuracy_vs_fp64"}
        diff = a32.astype(np.float64) - out64
        accuracy = {
            "reference": ("FP64 GPU run (fused kernel, within 1e-12 rel L2 of the CPU oracle "
                          "after 1000 steps in the tests) of the same steps from the same "
                          "FP32-rounded field"),
            "steps": CFG2_STEPS,
            "max_abs_err": float(np.abs(diff).max()),
            "rel_l2_err": float(np.linalg.norm(diff) / np.linalg.norm(out64)),
        }
        del a32, out64, diff
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

    extra = {}
    if not args.no_extra:
        for name, fn in extra_configs(torch, stream, args):
            t0 = time.perf_counter()
            try:
                with torch.cuda.stream(stream):
                    extra[name] = fn()
            except Exception as exc:  # reported, never hidden
                extra[name] = {"error": f"{type(exc).__name__}: {exc}"}
            print(f"[extra] {name} done in {time.perf_counter() - t0:.1f} s", file=sys.stderr,
                  flush=True)

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
                   "steps_per_launch": spl, "launch_shape": tuned},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "min_bytes_per_launch": min_bytes_per_launch,
                     "avg_launch_us": round(avg_launch_s * 1e6, 3),
                     "unfused_equivalent_GBs": round(unfused_equivalent, 1),
                     "unfused_equivalent_frac": round(unfused_equivalent / HBM_PEAK_GBS, 4),
                     "valu_per_site_update": valu,
                     "valu_issue_frac": (None if valu is None else
                                         round(value * 1e9 * valu / VALU_PEAK_LANE_OPS, 4)),
                     "fabric_bytes_per_site_update": (None if traffic is None
                                                      else traffic / (sites * spl)),
                     "fabric_GBs": (None if traffic is None
                                    else round(traffic / avg_launch_s / 1e9, 1)),
                     "note": ("achieved = the least HBM bytes a launch needs (one read and one "
                              "write of the field, 8 B x sites, whatever the steps per launch) / "
                              "average launch time (HIP events on the plan's stream). "
                              "unfused_equivalent = 8 B per site-update x sites x steps per "
                              "launch / launch time, the traffic the same steps would need "
                              "one launch each. traffic = calibrated FETCH_SIZE + WRITE_SIZE "
                              "bytes per launch of this shape (rocprofv3, profiles/); they count "
                              "L2-miss traffic to the fabric, which the Infinity Cache serves "
                              "for this 128 MiB working set, hence fabric_GBs. valu_issue_frac "
                              "= VALU lane-instructions per site-update (SQ_INSTS_VALU x 64, same "
                              "profile) x site-updates/s / 39.3e12")},
        "cpu_baseline": cpu,
        "accuracy_vs_fp64": accuracy,
        "fp64_rate": fp64,
        "mass_drift_rel": drift,
        "energy": {"initial": e0, "final": e1},
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
