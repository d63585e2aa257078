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


TRAFFIC_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                               "pmc_step3_kernel_d3q19_192x96x96_fp32.md")


def measured_traffic():
    """HBM bytes per launch of the FP32 step kernel on this workload, from the committed PMC summary
    (FETCH_SIZE and WRITE_SIZE passes calibrated against a copy kernel of known traffic)."""
    try:
        text = open(TRAFFIC_PROFILE).read()
        block = text.split("```json", 1)[1].split("```", 1)[0]
        return int(json.loads(block)["traffic_bytes_per_launch"])
    except (OSError, IndexError, KeyError, ValueError):
        return None


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


# Launch shapes pinned from the command line (--launch / --batch-launch) instead of autotuned, so that
# counter passes, which slow every dispatch and can change the autotuner's choice, measure one shape.
PIN = {"launch": None, "batch": None, "stride": None, "tiled": False}


def tune_or_pin(s, trial_steps):
    if PIN["launch"] is not None and s.Q == 19:
        s.set_launch(*PIN["launch"])
        block = s.launch_info()[0]
    else:
        block = s.autotune(trial_steps)
    if PIN["stride"] is not None and s.Q == 19:
        s.set_population_stride(PIN["stride"])
    if PIN["tiled"] and s.Q == 19:
        s.set_population_layout(True)
    return block


def run_gpu(W, H, D, precision, steps, warmup, profile=True, velocity_set=19):
    from cells_amd.lbm import Solver3D, bytes_per_cell

    s = Solver3D.create(dict(tunnel_config(W, H, D, precision), velocitySet=velocity_set))
    s.set_sdf(tunnel_sdf(W, H, D))
    s.init_field()
    s.set_inlet_ramp("smoothstep", 0.1, 3000)
    block = tune_or_pin(s, 200)
    if warmup:
        s.time_steps(warmup)
    s.sync()
    ms = s.time_steps(steps)
    s.sync()
    N = s.N
    mlups = N * steps / (ms * 1e3)
    out = dict(precision=precision, grid=[W, H, D], cells=N, steps=steps, ms=ms, ms_per_step=ms / steps,
               mlups=mlups, bytes_per_cell=bytes_per_cell(s.Q, s.precision), velocity_set=s.Q)
    out["algorithmic_gbs"] = mlups * 1e6 * out["bytes_per_cell"] / 1e9
    out["roofline_frac"] = out["algorithmic_gbs"] / HBM_PEAK_GBS
    nb, nl, ns = s.node_counts()
    out["nodes"] = dict(bulk=nb, list=nl, solid=ns)
    out["bulk_block"], out["row_order"] = s.launch_info()
    # N, the padded stride or the tiles of 64 cells the autotuner chose
    out["population_tiled"], out["population_stride"] = s.population_layout()
    # One kernel launch per step, back to back on the solver's stream: the step kernel's average
    # duration is the time between the HIP events around the timed steps over their count (a kernel
    # trace of the same region gives the same average, the launches leave no gaps).
    out["step_kernel_us"] = ms / steps * 1e3
    out["bytes_per_launch"] = N * out["bytes_per_cell"]
    out["kernel_achieved_gbs"] = out["bytes_per_launch"] / (ms / steps * 1e-3) / 1e9
    if profile:
        # an event after every step: the slowest single step (the extra events add gaps, so the
        # per-step mean here reads a little above the kernel's own duration)
        tot, tsum, tmax = s.profile_steps(steps)
        out["evented_step_us"] = tsum / steps * 1e3
        out["slowest_step_us"] = tmax * 1e3
    forces = s.read_forces()
    out["last_force"] = list(forces[-1][1:]) if forces else None
    out["finite"] = bool(np.isfinite(s.read_populations()).all())
    s.destroy()
    return out


def d2q9_config0(steps=5000):
    """D2Q9 512 x 128 cylinder tunnel (diameter 16 at (128, 64.5), tau 0.548, Zou-He inlet u = 0.1
    ramped linearly over 3000 steps, Zou-He outlet, slip walls, 15% sponge, absorbing layer 0.02,
    a force sample every step): MLUPS and algorithmic TB/s at 76 B per cell. At 65,536 cells one step
    is a few microseconds, so this measures launch-bound throughput; the floor is reported beside it:
    the copy-roofline kernel (9 populations and the flag word in, 9 out, no arithmetic) at the same
    size, one launch per step, and the same tunnel without force samples."""
    from cells_amd import geometry as g
    from cells_amd.lbm import X_OPEN, Y_SLIP, Solver, copy_roofline

    W, H = 512, 128

    def run(force_every):
        s = Solver.create(dict(width=W, height=H, left=X_OPEN, right=X_OPEN, bottom=Y_SLIP, top=Y_SLIP, tau=0.548,
                               uIn=0.1, spongeFraction=0.15, spongeTau=1, absorb=0.02, forceEvery=force_every))
        sdf = g.empty_sdf(W, H)
        g.add_circle(sdf, W, H, 128, 64.5, 8)
        s.set_sdf(sdf)
        s.init_field()
        s.set_inlet_ramp("linear", 0.1, 3000)
        block = tune_or_pin(s, 200)
        s.time_steps(500)
        ms = s.time_steps(steps)
        s.destroy()
        return ms, block

    ms, block = run(1)
    ms_nf, _ = run(0)
    mlups = W * H * steps / (ms * 1e3)
    copy_us = min(W * H * 76 / (copy_roofline(W * H, 9, 4, iters=2000, workgroup=wg) * 1e9) * 1e6 for wg in (64, 128, 256))
    return dict(grid=[W, H], steps=steps, mlups=mlups, ms_per_step=ms / steps, algorithmic_gbs=mlups * 76e-3,
                bulk_block=block, step_us=ms / steps * 1e3, step_without_force_samples_us=ms_nf / steps * 1e3,
                copy_kernel_us_same_size=copy_us, copy_kernel_mlups_same_size=W * H / copy_us)


PRECISION_FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden",
                                 "c1_sphere_192x96x96_fp64_forces.npy")
STEADY_FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden",
                              "c1_sphere_192x96x96_fp64_steady.npy")


def _window_errors(cd, cd64):
    return dict(cd_mean=float(cd.mean()), cd64_mean=float(cd64.mean()),
                rel_err_mean=float(abs(cd.mean() - cd64.mean()) / abs(cd64.mean())),
                max_abs_err_over_mean=float(np.abs(cd - cd64).max() / abs(cd64.mean())))


def precision_study(steps=20000):
    """C1 tunnel from the initial field in FP32, FP16 x 2^15 (rounded to nearest, and stochastically) and BF16
    storage: C_D (2 F_x / (u^2 pi 8^2))
    against the FP64 run of the same tunnel (the oracle in double precision) at identical steps, over
    two windows: steps 3001-4000, the first 1000 after the inlet ramp, while the wake still develops
    (tests/golden/make_c1_fp64_forces.py), and steps 19001-20000, where C_D is steady
    (tests/golden/make_c1_fp64_steady.py; needs steps >= 20000)."""
    from cells_amd.lbm import Solver3D

    W, H, D, u = 192, 96, 96, 0.1
    scale = 2 / (u * u * np.pi * 8 * 8)
    ref = np.load(PRECISION_FIXTURE) if os.path.exists(PRECISION_FIXTURE) else None
    steady = np.load(STEADY_FIXTURE) if os.path.exists(STEADY_FIXTURE) else None
    out = {}
    for prec in ("fp32", "fp16", "fp16sr", "bf16"):
        s = Solver3D.create(tunnel_config(W, H, D, prec))
        s.set_sdf(tunnel_sdf(W, H, D))
        s.init_field()
        s.set_inlet_ramp("smoothstep", u, 3000)
        sts, vals = [], []
        for k in range(0, steps, 4000):  # the force ring holds 8192 samples
            s.encode_steps(min(4000, steps - k))
            st, v = s.read_forces_raw()
            sts.append(st)
            vals.append(v)
        s.destroy()
        st = np.concatenate(sts)
        cd = scale * np.concatenate(vals).reshape(-1, 3)[:, 0]
        assert np.array_equal(st, np.arange(1, steps + 1)), "one sample per step expected"
        r = dict(steps=steps, cd_last=float(cd[-1]))
        if ref is not None and steps >= 4000:
            r["transient_3001_4000"] = _window_errors(cd[3000:4000], scale * ref[3000:4000, 0])
        if steady is not None and steps >= 20000:
            r["steady_19001_20000"] = _window_errors(cd[19000:20000], scale * steady[:, 0])
        out[prec] = r
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
    """The 64-tunnel NACA0012 sweep: one batch (autotuned launch shape) vs the same tunnels as
    sequential plain runs, compared with divergence taken into account (cells_amd.sweep.compare_runs)."""
    from cells_amd import sweep

    cases = sweep.sweep_cases(64, seed=0)
    ib, iseq = {}, {}
    tb, rb = sweep.run_batched(cases, steps, info=ib, launch=PIN["batch"])
    ts, rs = sweep.run_sequential(cases, steps, info=iseq)
    cmp = sweep.compare_runs(cases, rb, rs, min(2000, steps // 2))
    # the device-side divergence records agree with the traces and between the two runs
    cmp["device_divergence_records_agree"] = bool(ib["diverged_at"] == iseq["diverged_at"] and all(
        d["first_bad_step"] == ib["diverged_at"][d["tunnel"]] for d in cmp["diverged"]) and sum(
        v != 0 for v in ib["diverged_at"]) == len(cmp["diverged"]))
    bm, sm = sweep.mlups(cases, steps, tb), sweep.mlups(cases, steps, ts)
    cells = sum(c["config"]["width"] * c["config"]["height"] for c in cases)
    return dict(cells=cells, steps=steps, batched_mlups=bm, sequential_mlups=sm, speedup=bm / sm,
                batched_s=tb, sequential_s=ts, launch=dict(tile=ib["tile"], order=ib["order"]),
                batched_algorithmic_gbs=bm * 76e-3, batched_roofline_frac=bm * 76e-3 / HBM_PEAK_GBS, **cmp)


def adjoint2d_case(fd=True):
    """2D shape gradient: D2Q9 FP32 384 x 128, NACA0012 of chord 64 at 5 degrees (leading edge at
    (96, 63.5)), Re 400 (tau 0.548), Zou-He inlet u = 0.1 ramped linearly over 3000 steps, Zou-He
    outlet, slip walls, 15% sponge, absorbing layer 0.02; J = mean C_D over steps 4001..6000;
    parameters (alpha, a_1..a_8) with bump width chord / 16; checkpoints every 200 steps."""
    from cells_amd import shapes
    from cells_amd.adjoint import ShapeAdjoint, bouzidi_branch_switches, parameter_gradient
    from cells_amd.lbm import X_OPEN, Y_SLIP, Solver

    W, H, chord, u, T, window, K = 384, 128, 64.0, 0.1, 6000, 2000, 200
    tau = 3 * u * chord / 400 + 0.5
    cfg = dict(width=W, height=H, left=X_OPEN, right=X_OPEN, bottom=Y_SLIP, top=Y_SLIP, tau=tau, uIn=u,
               spongeFraction=0.15, spongeTau=1, absorb=0.02, forceEvery=1)
    theta = np.array([5.0] + [0.0] * 8)

    def sdf_of(th):
        return shapes.naca_bumps(W, H, 96.0, 63.5, chord, th[0], th[1:])

    s = Solver.create(dict(cfg))
    sdf0, dsdf = sdf_of(theta)
    s.set_sdf(sdf0)

    def plain_forward():
        # the same run through the ordinary step path
        s.set_inlet_ramp("linear", u, 3000)
        s.init_field()
        s.sync()
        t0 = time.perf_counter()
        s.encode_steps(T)
        s.sync()
        return time.perf_counter() - t0

    def timed(budget):
        """(adjoint forward s, backward s, J, dJ/dsdf, info): the fastest of 3 rounds after one untimed
        run that builds the reverse-time tables."""
        adj = ShapeAdjoint(s, T, window, (1.0, 0.0), 2 / (u * u * chord), checkpoint_every=K,
                           ramp=("linear", u, 3000), record_budget_bytes=budget)
        adj.forward()
        adj.info()
        tf, tb = [], []
        for _ in range(3):
            t0 = time.perf_counter()
            J = adj.forward()
            t1 = time.perf_counter()
            g_sdf = adj.backward()
            tf.append(t1 - t0)
            tb.append(time.perf_counter() - t1)
        return min(tf), min(tb), J, g_sdf, adj.info(), adj

    t_plain = min(plain_forward() for _ in range(3))
    # the configured plan: a checkpoint every 200 steps, every segment recomputed in the backward pass
    tf0, tb0, J, g_sdf, info, adj = timed(0)
    # every segment recorded by the forward pass instead (16 B per cell and step; 4.7 GB here)
    tf1, tb1, J1, g1, info1, _ = timed(16 << 30)
    assert J1 == J and np.array_equal(g1, g_sdf), "recorded and recomputed segments must agree bit for bit"
    grad = parameter_gradient(g_sdf, dsdf)
    out = dict(grid=[W, H], steps=T, J=J, grad=grad.tolist(), info=info, plain_forward_s=t_plain,
               adjoint_forward_s=tf0, backward_s=tb0, ratio=(tf0 + tb0) / t_plain,
               backward_over_forward=tb0 / t_plain, memory_plan="checkpoint every 200 steps, recompute",
               recorded=dict(info=info1, adjoint_forward_s=tf1, backward_s=tb1, ratio=(tf1 + tb1) / t_plain),
               timing="fastest of 3 rounds")
    # the FP64 reverse-mode reference of the same case (committed fixture, CPU torch in float64)
    try:
        ref = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden",
                                          "adjoint_naca_384x128_fp64.json")))
        g_ref = np.array(ref["grad_theta"])
        out["J_fp64"] = ref["J"]
        out["max_rel_err_vs_fp64"] = float(np.max(np.abs(grad - g_ref) / np.abs(g_ref)))
    except (OSError, KeyError, ValueError):
        pass
    try:  # central differences of the FP64 oracle (tests/golden/make_adjoint_fd_fp64.py)
        ref = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden",
                                          "adjoint_fd_fp64.json")))["naca_384x128"]
        for h, key in (("1e-3", "fd_h0.001"), ("1e-4", "fd_h0.0001")):
            fdv = np.array(ref[key])
            out[f"max_rel_err_vs_fp64_fd_h{h}"] = float(np.max(np.abs(grad - fdv) / np.abs(fdv)))
        out["fp64_fd_solid_flips"] = int(sum(ref["flips_h0.001"]) + sum(ref["flips_h0.0001"]))
    except (OSError, KeyError, ValueError):
        pass
    if fd:
        # central differences of the GPU's FP32 forward at h = 1e-3 and 1e-2 (at 1e-3 the smallest
        # components are within a few FP32 roundings of J); flips counts the nodes whose solid flag
        # differs between theta +- h e_p and theta, which would make J jump, and branch_switches per
        # parameter the Bouzidi links whose q lies on different sides of 1/2 (or of a clamp bound) at
        # theta + h e_p and theta - h e_p: J's slope changes there, so a difference quotient across one
        # is not J's derivative at theta. The error over the parameters without a switch is reported
        # beside the error over all of them.
        solid0 = sdf0.reshape(H, W)[:, 1:-1] < 0
        for h, key in ((1e-3, "fd"), (1e-2, "fd_h1e-2")):
            fdv, flips, switches = [], 0, []
            for p in range(len(theta)):
                e = np.zeros_like(theta)
                e[p] = h
                sp, sm = sdf_of(theta + e)[0], sdf_of(theta - e)[0]
                flips += int(np.count_nonzero((sp.reshape(H, W)[:, 1:-1] < 0) != solid0))
                flips += int(np.count_nonzero((sm.reshape(H, W)[:, 1:-1] < 0) != solid0))
                switches.append(bouzidi_branch_switches(sp, sm, W, H))
                fdv.append((adj.forward(sp) - adj.forward(sm)) / (2 * h))
            fdv = np.array(fdv)
            out[key] = fdv.tolist()
            out[key + "_solid_flips"] = flips
            out[key + "_branch_switches"] = switches
            out["max_rel_err_vs_" + key] = float(np.max(np.abs(grad - fdv) / np.abs(fdv)))
            out["max_err_over_max_grad_vs_" + key] = float(np.max(np.abs(grad - fdv)) / np.abs(fdv).max())
            smooth = [p for p in range(len(theta)) if switches[p] == 0]
            if smooth:
                out["max_rel_err_vs_" + key + "_without_switches"] = float(
                    np.max(np.abs(grad[smooth] - fdv[smooth]) / np.abs(fdv[smooth])))
    s.destroy()
    return out


def adjoint3d_case(fd_coeffs=(0, 1, 2, 3, 6, 12), budget_gb=(210, 120, 0)):
    """3D shape gradient: D3Q19 256 x 128 x 128, sphere of diameter 24 at (64, 63.5, 63.5) (off the
    lattice in y and z, so that no node lies on the surface and a small radius change flips no node
    between fluid and solid) with its radius
    perturbed by 16 real spherical harmonics (l <= 3, all 0), Re 300 (tau 0.524), equilibrium inlet
    u = 0.1 ramped (smoothstep) over 1000 steps, slip walls, 15% sponge, absorbing layer 0.02.
    Forward in FP16 x 2^15 storage, adjoint in FP32 with FP64 dJ/dq; J = mean C_D over steps
    2001..3000 (area pi r^2); checkpoints every 100 steps, the last segments recorded directly within
    the device-memory budget (the first of the budgets that can be allocated). Finite differences
    (h = 1e-3) use the FP32-storage forward, whose own adjoint gradient is reported alongside."""
    from cells_amd import shapes
    from cells_amd.adjoint import ShapeAdjoint, bouzidi_branch_switches, parameter_gradient
    from cells_amd.lbm import Solver3D

    W, H, D, u, T, window, K, r = 256, 128, 128, 0.1, 3000, 1000, 100, 12.0
    tau = 3 * u * 2 * r / 300 + 0.5
    area = np.pi * r * r
    coeffs = np.zeros(16)

    def sdf_of(b):
        return shapes.sphere_sh(W, H, D, (64.0, 63.5, 63.5), r, b)

    def cfg(precision):
        return dict(width=W, height=H, depth=D, precision=precision, tau=tau, uIn=u, spongeFraction=0.15,
                    spongeTau=1, absorb=0.02, forceEvery=1)

    sdf0, dsdf = sdf_of(coeffs)
    s = Solver3D.create(cfg("fp16"))
    s.set_sdf(sdf0)
    tune_or_pin(s, 20)
    s.set_inlet_ramp("smoothstep", u, 1000)

    def plain_forward():
        s.init_field()
        s.sync()
        t0 = time.perf_counter()
        s.encode_steps(T)
        s.sync()
        return time.perf_counter() - t0

    from cells_amd.lbm import LbmError

    def timed(budgets):
        """The first budget that can be allocated: (budget GB, forward s, backward s, J, dJ/dsdf, info),
        the fastest of 3 rounds after one untimed run that builds the reverse-time tables."""
        for gb in budgets:
            adj = ShapeAdjoint(s, T, window, (1.0, 0.0, 0.0), 2 / (u * u * area), checkpoint_every=K,
                               ramp=("smoothstep", u, 1000), record_budget_bytes=int(gb * (1 << 30)))
            try:
                adj.forward()
                break
            except LbmError:
                adj.release()
        info = adj.info()
        tf, tb = [], []
        for _ in range(3):
            t0 = time.perf_counter()
            J = adj.forward()
            t1 = time.perf_counter()
            g_sdf = adj.backward()
            tf.append(t1 - t0)
            tb.append(time.perf_counter() - t1)
        adj.release()
        return gb, min(tf), min(tb), J, g_sdf, info

    t_plain = min(plain_forward() for _ in range(3))
    # the configured plan: a checkpoint every 100 steps (30 x 159 MB in FP16), every segment
    # recomputed in the backward pass into one recorded segment's worth of scratch (6.7 GB)
    _, t_fwd, t_bwd, J, g_sdf, info = timed((0,))
    # every segment recorded by the forward pass instead (8 B per cell and step in FP16 storage, 101 GB)
    gb, tf1, tb1, J1, g1, info1 = timed(budget_gb)
    grad = parameter_gradient(g_sdf, dsdf)
    s.destroy()
    out = dict(grid=[W, H, D], steps=T, J=J, grad=grad.tolist(), info=info, plain_forward_s=t_plain,
               adjoint_forward_s=t_fwd, backward_s=t_bwd, ratio=(t_fwd + t_bwd) / t_plain,
               backward_over_forward=t_bwd / t_plain, memory_plan="checkpoint every 100 steps, recompute",
               recorded=dict(record_budget_gb=gb, info=info1, adjoint_forward_s=tf1, backward_s=tb1,
                             ratio=(tf1 + tb1) / t_plain,
                             same_gradient_bits=bool(J1 == J and np.array_equal(g1, g_sdf))),
               timing="fastest of 3 rounds")
    if fd_coeffs:
        s32 = Solver3D.create(cfg("fp32"))
        a32 = ShapeAdjoint(s32, T, window, (1.0, 0.0, 0.0), 2 / (u * u * area), checkpoint_every=K,
                           ramp=("smoothstep", u, 1000), record_budget_bytes=64 << 30)
        out["J_fp32"] = a32.forward(sdf0)
        g32 = parameter_gradient(a32.backward(), dsdf)
        out["grad_fp32"] = g32.tolist()
        a32.release()
        h = 1e-3
        fdv, switches = {}, {}
        for p in fd_coeffs:
            e = np.zeros(16)
            e[p] = h
            sp, sm = sdf_of(coeffs + e)[0], sdf_of(coeffs - e)[0]
            switches[str(p)] = bouzidi_branch_switches(sp, sm, W, H, D)
            fdv[p] = (a32.forward(sp) - a32.forward(sm)) / (2 * h)
        s32.destroy()
        scale = max(abs(v) for v in fdv.values())
        out["fd_fp32"] = {str(p): v for p, v in fdv.items()}
        # Bouzidi links whose q changes side of 1/2 (or of a clamp bound) between coeffs +- h e_p
        out["fd_branch_switches"] = switches
        # relative error over the components above 1% of the largest (the others are zero by the
        # body's symmetry, where differences only see rounding), and the largest error over all of
        # them relative to the largest component
        big = [p for p, v in fdv.items() if abs(v) >= 0.01 * scale]
        for name, g in (("", grad), ("fp32_adjoint_", g32)):
            out[f"max_rel_err_{name}vs_fd"] = float(max(abs(g[p] - fdv[p]) / abs(fdv[p]) for p in big))
            out[f"max_err_over_max_grad_{name}vs_fd"] = float(max(abs(g[p] - v) for p, v in fdv.items()) / scale)
    return out


# Single workloads for profiling (bench.py --only NAME): the D3Q19 / D3Q27 tunnels of the bench line
# at --steps steps, the 64-tunnel sweep at --sweep-steps, the D2Q9 512 x 128 cylinder, the copy
# roofline kernels, and the two shape-gradient cases without finite differences.
ONLY = {
    "fp32_192x96x96": lambda a: run_gpu(192, 96, 96, "fp32", a.steps, a.warmup, profile=False),
    "fp16_192x96x96": lambda a: run_gpu(192, 96, 96, "fp16", a.steps, a.warmup, profile=False),
    "fp16sr_192x96x96": lambda a: run_gpu(192, 96, 96, "fp16sr", a.steps, a.warmup, profile=False),
    "bf16_192x96x96": lambda a: run_gpu(192, 96, 96, "bf16", a.steps, a.warmup, profile=False),
    "fp32_256x128x128": lambda a: run_gpu(256, 128, 128, "fp32", a.steps, a.warmup, profile=False),
    "fp16_256x128x128": lambda a: run_gpu(256, 128, 128, "fp16", a.steps, a.warmup, profile=False),
    "d3q27_fp32_192x96x96": lambda a: run_gpu(192, 96, 96, "fp32", a.steps, a.warmup, profile=False, velocity_set=27),
    "d3q27_fp16_192x96x96": lambda a: run_gpu(192, 96, 96, "fp16", a.steps, a.warmup, profile=False, velocity_set=27),
    "d2q9_512x128_cylinder": lambda a: d2q9_config0(a.steps),
    "sweep_64": lambda a: sweep_bench(a.sweep_steps),
    "adjoint_naca_384x128": lambda a: adjoint2d_case(fd=False),
    "adjoint_sphere_256x128x128": lambda a: adjoint3d_case(fd_coeffs=()),
    "copy_roofline": lambda a: copy_rooflines(),
}


def filled_pull_key(variant):
    """The filled pull-copy entry of copy_rooflines() with the element width and grid of a variant."""
    prec, size = variant.split("_", 1)
    return f"filled_pull_{'fp32' if prec == 'fp32' else 'fp16'}_{size}_gbs"


def copy_rooflines():
    """Copy-roofline kernels (src/bench.ts:61-95: Q populations and the flag word read, Q written, no
    arithmetic, aligned streams) and the pull copy (the same bytes with a D3Q19 step's shifted streams).
    The reference page copies zeroed buffers; the memory system moves zeros faster than the values a
    step moves, so each is also measured over flow-like values ("filled"), the ceiling for a step."""
    from cells_amd.lbm import copy_roofline, pull_copy_roofline

    n1, n4 = 192 * 96 * 96, 256 * 128 * 128
    out = {}
    for tag, filled in (("", False), ("filled_", True)):
        out.update({f"{tag}fp32_192x96x96_gbs": copy_roofline(n1, 19, 4, filled=filled),
                    f"{tag}fp16_192x96x96_gbs": copy_roofline(n1, 19, 2, filled=filled),
                    f"{tag}fp32_256x128x128_gbs": copy_roofline(n4, 19, 4, filled=filled),
                    f"{tag}fp16_256x128x128_gbs": copy_roofline(n4, 19, 2, filled=filled),
                    f"{tag}d2q9_fp32_7430144_gbs": copy_roofline(7430144, 9, 4, filled=filled),
                    f"{tag}pull_fp32_192x96x96_gbs": pull_copy_roofline(192, 96, 96, 4, filled=filled),
                    f"{tag}pull_fp16_192x96x96_gbs": pull_copy_roofline(192, 96, 96, 2, filled=filled),
                    f"{tag}pull_fp32_256x128x128_gbs": pull_copy_roofline(256, 128, 128, 4, filled=filled),
                    f"{tag}pull_fp16_256x128x128_gbs": pull_copy_roofline(256, 128, 128, 2, filled=filled)})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="FP32 192x96x96 only (for profiling)")
    ap.add_argument("--sweep-steps", type=int, default=10000, help="steps of the 64-tunnel sweep (0 = skip)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the 2D and 3D shape-gradient runs")
    ap.add_argument("--only", choices=sorted(ONLY), help="run one workload and print its own JSON line "
                    "(for kernel traces and counter passes of that workload)")
    ap.add_argument("--launch", help="pin the D2Q9 / D3Q19 launch shape: BLOCK,VARIANT (lbm_set_launch)")
    ap.add_argument("--batch-launch", help="pin the batch launch shape: TILE,ORDER (lbm_batch_set_launch)")
    ap.add_argument("--stride", type=int, help="pin the D3Q19 population stride in elements, 0 = N "
                    "(lbm_set_population_stride)")
    ap.add_argument("--tiled", action="store_true", help="pin the D3Q19 tiled population layout "
                    "(lbm_set_population_layout)")
    args = ap.parse_args()
    if args.launch:
        PIN["launch"] = tuple(int(v) for v in args.launch.split(","))
    if args.batch_launch:
        PIN["batch"] = tuple(int(v) for v in args.batch_launch.split(","))
    if args.stride is not None:
        PIN["stride"] = args.stride
    PIN["tiled"] = args.tiled
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the solver does not shard "
              f"across GPUs; see the sharding design notes).", file=sys.stderr)
        sys.exit(2)
    if args.only:
        print(json.dumps({"only": args.only, "result": ONLY[args.only](args)}))
        return

    W, H, D = 192, 96, 96
    main_run = run_gpu(W, H, D, "fp32", args.steps, args.warmup)
    variants = {"fp32_192x96x96": main_run}
    if not args.quick:
        for prec in ("fp16", "fp16sr", "bf16"):
            variants[f"{prec}_192x96x96"] = run_gpu(W, H, D, prec, args.steps, args.warmup)
        for prec in ("fp32", "fp16"):
            variants[f"{prec}_256x128x128"] = run_gpu(256, 128, 128, prec, max(1, args.steps // 2),
                                                      args.warmup // 2)
        for prec in ("fp32", "fp16"):  # the D3Q27 velocity set on the same tunnel
            variants[f"d3q27_{prec}_192x96x96"] = run_gpu(W, H, D, prec, args.steps, args.warmup, velocity_set=27)
    copy = copy_rooflines()
    sweep_res = sweep_bench(args.sweep_steps) if (args.sweep_steps > 0 and not args.quick) else None
    precision = None if args.quick else precision_study()
    d2q9 = None if args.quick else d2q9_config0()
    adjoint = None
    if not (args.quick or args.no_adjoint):
        adjoint = {"naca_384x128": adjoint2d_case(), "sphere_256x128x128_fp16": adjoint3d_case()}
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
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     "kernel": "step3_kernel<StoreF32>",
                     "bytes_per_launch": main_run["bytes_per_launch"],
                     "avg_launch_us": main_run["step_kernel_us"]},
        "hbm_roofline_frac_whole_step": main_run["roofline_frac"],
        "cpu_baseline": cpu,
        "variants": {k: {kk: v[kk] for kk in ("mlups", "ms_per_step", "algorithmic_gbs", "roofline_frac",
                                               "step_kernel_us", "evented_step_us", "slowest_step_us",
                                               "kernel_achieved_gbs", "bytes_per_cell",
                                               "bulk_block", "row_order", "population_tiled", "population_stride",
                                               "nodes", "finite") if kk in v}
                     for k, v in variants.items()},
        "fp16_over_fp32": variants.get("fp16_192x96x96", main_run)["mlups"] / main_run["mlups"],
        "fp16sr_over_fp32": variants.get("fp16sr_192x96x96", main_run)["mlups"] / main_run["mlups"],
        "bulk_block": main_run["bulk_block"],
        "copy_roofline": copy,
        # each step against the pull copy of its size and element width over flow-like values (the
        # same bytes and shifted streams with no arithmetic)
        "step_over_filled_pull_copy": {
            k: variants[k]["algorithmic_gbs"] / copy[filled_pull_key(k)]
            for k in ("fp32_192x96x96", "fp16_192x96x96", "fp16sr_192x96x96", "bf16_192x96x96", "fp32_256x128x128",
                      "fp16_256x128x128") if k in variants},
        "sweep": sweep_res,
        "precision_vs_fp64": precision,
        "d2q9_512x128_cylinder": d2q9,
        "adjoint": adjoint,
        "adjoint_over_forward": adjoint["sphere_256x128x128_fp16"]["ratio"] if adjoint else None,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
