#!/usr/bin/env python3
"""Throughput benchmark of the D3Q19 step on one MI355X.

Workload: D3Q19 TRT tunnel, 192 x 96 x 96 (1,769,472 cells), sphere of diameter 16 at (48, 48, 48)
with Bouzidi walls, equilibrium inlet u = 0.1 ramped over 3000 steps, pressure outlet, 15% sponge,
absorbing layer 0.02, slip side faces, Re = 200 (tau = 0.524, Lambda = min(3/16, 50 (tau - 1/2)^2)),
one force sample per step. Inputs are synthetic and resident in HBM before timing starts.

One "step" is one time step of the whole tunnel: one launch of the step kernel (bulk rows plus the
boundary-list blocks with the FP64 force reduction). value = cells * steps / time in MLUPS for FP32
storage; FP16 x 2^15 and BF16 storage, and the same tunnel at 256 x 128 x 128, are reported
alongside. The roofline object describes that kernel from HIP events after every launch on the
solver's stream, at the algorithmic 156 B per cell (19 populations read and written in FP32 plus the
flag word) over all cells; the CPU baseline is the oracle (the FP32 C++ restatement) timed on a
bounded sample of the same workload on the host cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
REF_M5_FP32_MLUPS = 708.0  # reference, D3Q19 192x96x96 FP32 on Apple M5 (docs/dietpowers/2026-09-27-3d-tunnel-spec.md:291)


def tunnel_config(W, H, D, precision):
    from cells_amd.analysis import derive_tau

    tau = derive_tau(200, 0.1, 16).tau
    return dict(width=W, height=H, depth=D, precision=precision, tau=tau, uIn=0.0, spongeFraction=0.15,
                spongeTau=1, absorb=0.02, forceEvery=1, workgroupSize=256)


def tunnel_sdf(W, H, D):
    from cells_amd import geometry as g

    sdf = g.empty_sdf3(W, H, D)
    g.add_sphere(sdf, W, H, D, W / 4, H / 2, D / 2, 8)
    return sdf


def run_gpu(W, H, D, precision, steps, warmup, profile=True):
    from cells_amd.lbm import Solver3D, bytes_per_cell

    s = Solver3D.create(tunnel_config(W, H, D, precision))
    s.set_sdf(tunnel_sdf(W, H, D))
    s.init_field()
    s.set_inlet_ramp("smoothstep", 0.1, 3000)
    block = s.autotune(40)
    if warmup:
        s.time_steps(warmup)
    s.sync()
    ms = s.time_steps(steps)
    s.sync()
    N = s.N
    mlups = N * steps / (ms * 1e3)
    out = dict(precision=precision, grid=[W, H, D], cells=N, steps=steps, ms=ms, ms_per_step=ms / steps,
               mlups=mlups, bytes_per_cell=bytes_per_cell(19, s.precision))
    out["algorithmic_gbs"] = mlups * 1e6 * out["bytes_per_cell"] / 1e9
    out["roofline_frac"] = out["algorithmic_gbs"] / HBM_PEAK_GBS
    nb, nl, ns = s.node_counts()
    out["nodes"] = dict(bulk=nb, list=nl, solid=ns)
    out["bulk_block"] = block
    if profile:
        # one kernel launch per step: HIP events after every step on the solver's stream
        tot, tsum, tmax = s.profile_steps(steps)
        out["step_kernel_us"] = tsum / steps * 1e3
        out["slowest_step_us"] = tmax * 1e3
        out["bytes_per_launch"] = N * out["bytes_per_cell"]
        out["kernel_achieved_gbs"] = out["bytes_per_launch"] / (tsum / steps * 1e-3) / 1e9
    forces = s.read_forces()
    out["last_force"] = list(forces[-1][1:]) if forces else None
    out["finite"] = bool(np.isfinite(s.read_populations()).all())
    s.destroy()
    return out


def cpu_baseline(W, H, D, budget_s=12.0):
    """The oracle (FP32 C++ restatement, OpenMP) on the same tunnel for a bounded number of steps."""
    from oracle import oracle

    lib = oracle.load()
    threads = int(lib.orc_max_threads())
    s = oracle.OracleSolver3D(tunnel_config(W, H, D, "fp32"), 32)
    s.set_sdf(tunnel_sdf(W, H, D))
    s.init_field()
    s.set_inlet_ramp("smoothstep", 0.1, 3000)
    s.steps(2)  # warm-up
    t0 = time.perf_counter()
    s.steps(2)
    per = (time.perf_counter() - t0) / 2
    n = max(2, int(budget_s / max(per, 1e-6)))
    t0 = time.perf_counter()
    s.steps(n)
    dt = time.perf_counter() - t0
    N = W * H * D
    cpu = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return dict(value=N * n / dt / 1e6, unit="MLUPS", cores=threads, kind="port",
                sample=f"oracle FP32 D3Q19 {W}x{H}x{D} sphere tunnel, {n} steps after 4 warm-up steps, "
                       f"{dt:.1f} s on {threads} threads of {cpu}")


def sweep_bench(steps):
    """The 64-tunnel NACA0012 sweep: one batch vs the same tunnels as sequential plain runs."""
    import numpy as np

    from cells_amd import sweep

    cases = sweep.sweep_cases(64, seed=0)
    tb, rb = sweep.run_batched(cases, steps)
    ts, rs = sweep.run_sequential(cases, steps)
    worst = 0.0
    for c, (sb, vb), (ss, vs) in zip(cases, rb, rs):
        cd_b, cl_b = sweep.coefficients(sb, vb, c["chord"], min(2000, steps // 2))
        cd_s, cl_s = sweep.coefficients(ss, vs, c["chord"], min(2000, steps // 2))
        worst = max(worst, abs(cd_b - cd_s) / max(abs(cd_s), 1e-12), abs(cl_b - cl_s) / max(abs(cl_s), 1e-3))
    bad = []
    for i, (c, (sb, vb)) in enumerate(zip(cases, rb)):
        ok = np.isfinite(vb).all(axis=1)
        if not ok.all():
            bad.append(dict(tunnel=i, nx=c["config"]["width"], alpha_deg=round(c["alpha_deg"], 3),
                            re=round(c["re"], 1), tau=c["config"]["tau"], first_bad_step=int(sb[~ok][0])))
    bm, sm = sweep.mlups(cases, steps, tb), sweep.mlups(cases, steps, ts)
    return dict(tunnels=len(cases), cells=sum(c["config"]["width"] * c["config"]["height"] for c in cases),
                steps=steps, batched_mlups=bm, sequential_mlups=sm, speedup=bm / sm,
                batched_s=tb, sequential_s=ts, mean_cd_cl_max_rel_diff=worst,
                finite=not bad, nonfinite_tunnels=bad)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="FP32 192x96x96 only (for profiling)")
    ap.add_argument("--sweep-steps", type=int, default=10000, help="steps of the 64-tunnel sweep (0 = skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the solver does not shard "
              f"across GPUs; see the sharding design notes).", file=sys.stderr)
        sys.exit(2)

    W, H, D = 192, 96, 96
    main_run = run_gpu(W, H, D, "fp32", args.steps, args.warmup)
    variants = {"fp32_192x96x96": main_run}
    if not args.quick:
        for prec in ("fp16", "bf16"):
            variants[f"{prec}_192x96x96"] = run_gpu(W, H, D, prec, args.steps, args.warmup)
        for prec in ("fp32", "fp16"):
            variants[f"{prec}_256x128x128"] = run_gpu(256, 128, 128, prec, max(1, args.steps // 2),
                                                      args.warmup // 2)
    from cells_amd.lbm import copy_roofline

    copy = {"fp32_192x96x96_gbs": copy_roofline(W * H * D, 19, 4), "fp16_192x96x96_gbs": copy_roofline(W * H * D, 19, 2),
            "fp32_256x128x128_gbs": copy_roofline(256 * 128 * 128, 19, 4)}
    sweep_res = sweep_bench(args.sweep_steps) if (args.sweep_steps > 0 and not args.quick) else None
    cpu = None if args.no_cpu else cpu_baseline(W, H, D)

    achieved = main_run["kernel_achieved_gbs"]
    line = {
        "metric": "D3Q19 MLUPS + % HBM roofline; batched sweep MLUPS; adjoint/forward time ratio",
        "value": main_run["mlups"],
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_run["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": main_run["mlups"] / REF_M5_FP32_MLUPS,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "D3Q19 TRT sphere tunnel 192x96x96, FP32 storage, Re 200, Bouzidi walls, "
                               "equilibrium inlet/outlet, sponge 15%, force sample every step",
                   "grid": [W, H, D], "cells": W * H * D, "bytes_per_cell": 156},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "step3_kernel<StoreF32>",
                     "bytes_per_launch": main_run["bytes_per_launch"],
                     "avg_launch_us": main_run["step_kernel_us"]},
        "hbm_roofline_frac_whole_step": main_run["roofline_frac"],
        "cpu_baseline": cpu,
        "variants": {k: {kk: v[kk] for kk in ("mlups", "ms_per_step", "algorithmic_gbs", "roofline_frac",
                                               "step_kernel_us", "slowest_step_us", "kernel_achieved_gbs",
                                               "bulk_block", "nodes", "finite") if kk in v}
                     for k, v in variants.items()},
        "fp16_over_fp32": variants.get("fp16_192x96x96", main_run)["mlups"] / main_run["mlups"],
        "bulk_block": main_run["bulk_block"],
        "copy_roofline": copy,
        "sweep": sweep_res,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
