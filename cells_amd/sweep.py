"""Batched sweep of NACA0012 tunnels (the multi-tunnel workload; no counterpart in the reference).

The draw, from numpy's PCG64 generator seeded with 0, per tunnel in this order:
  1. nx uniformly from {256, 384, 512, 768, 1024}; ny = nx / 4;
  2. angle of attack uniformly in [-10, 15] degrees;
  3. log10(Re) uniformly in [log10 100, log10 2000].
Each tunnel: NACA0012 of chord nx / 8 with its leading edge at (nx / 4, ny / 2 - 0.5), pitched
nose-up by the angle (src/geometry.ts:42-57 polygon, Bouzidi walls from its SDF); Zou-He inlet
u = 0.1 (smoothstep ramp over 3000 steps, src/cases.ts:52-53), Zou-He outlet, slip top/bottom, 15%
sponge with spongeTau 1 and absorbing layer 0.02 (src/app.ts:176-190 defaults); FP32 storage;
tau = 3 u chord / Re + 1/2 clamped at 0.51 with Smagorinsky C_s = 0.1 where the clamp acts (the
reference clamps tau in deriveTau, src/units.ts:14-20, and has one global C_s).
"""
from __future__ import annotations

import hashlib
import math

import numpy as np

from . import geometry as g
from .lbm import X_OPEN, Y_SLIP, Batch, Solver

U = 0.1
SIZES = (256, 384, 512, 768, 1024)


def sweep_cases(n: int = 64, seed: int = 0, sizes=SIZES):
    rng = np.random.default_rng(seed)
    cases = []
    for _ in range(n):
        nx = int(rng.choice(sizes))
        alpha = float(rng.uniform(-10.0, 15.0))
        re = float(10 ** rng.uniform(math.log10(100), math.log10(2000)))
        ny = nx // 4
        chord = nx / 8
        raw = 3 * U * chord / re + 0.5
        tau = max(raw, 0.51)
        cfg = dict(width=nx, height=ny, left=X_OPEN, right=X_OPEN, bottom=Y_SLIP, top=Y_SLIP, tau=tau, uIn=U,
                   spongeFraction=0.15, spongeTau=1, absorb=0.02, smagorinsky=0.1 if raw < 0.51 else 0.0,
                   forceEvery=1)
        cases.append(dict(config=cfg, alpha_deg=alpha, re=re, chord=chord))
    return cases


def case_sdf(case) -> np.ndarray:
    c = case["config"]
    W, H = c["width"], c["height"]
    sdf = g.empty_sdf(W, H)
    g.add_polygon(sdf, W, H, g.naca_polygon(W / 4, H / 2 - 0.5, case["chord"], 0.12, math.radians(case["alpha_deg"])))
    return sdf


def coefficients(steps, vals, chord, window):
    """Mean C_D and C_L over the samples with step > steps[-1] - window (C = 2 F / (u^2 chord))."""
    sel = steps > steps[-1] - window
    c = 2 * vals[sel] / (U * U * chord)
    return float(c[:, 0].mean()), float(c[:, 1].mean())


def _build_batch(cases, ramp_steps):
    b = Batch.create([c["config"] for c in cases])
    for i, c in enumerate(cases):
        t = b.tunnel(i)
        t.set_sdf(case_sdf(c))
        t.init_field()
    b.reset_clock()
    b.set_inlet_ramp("smoothstep", ramp_steps)
    return b


def run_batched(cases, steps: int, chunk: int = 4096, ramp_steps: float = 3000, autotune: bool = True,
                launch=None, info: dict = None):
    """All tunnels in one batch; returns (seconds of GPU stepping, per-tunnel (steps, forces)). The
    launch shape is autotuned on the initial state (or pinned by launch = (tile, order)); info, when
    given, receives the shape and per tunnel the step of its first non-finite force sample (0: none)."""
    b = _build_batch(cases, ramp_steps)
    if launch is not None:
        b.set_launch(*launch)
    elif autotune:
        b.autotune(40)
    traces = [([], []) for _ in cases]
    elapsed = 0.0
    done = 0
    while done < steps:
        k = min(chunk, steps - done)
        elapsed += b.time_steps(k) / 1e3
        for i in range(len(cases)):
            s, v = b.tunnels[i].read_forces_raw()
            traces[i][0].append(s)
            traces[i][1].append(v)
        done += k
    out = [(np.concatenate(s), np.concatenate(v)) for s, v in traces]
    if info is not None:
        info["tile"], info["order"] = b.launch_info()
        info["diverged_at"] = [int(v) for v in b.diverged_at()]
        info["checksums"] = [_digest(t) for t in b.tunnels]
    b.destroy()
    return elapsed, out


def _digest(solver) -> str:
    """SHA-1 of a tunnel's current population words (bit-for-bit comparisons without keeping them)."""
    return hashlib.sha1(solver.read_populations_raw().tobytes()).hexdigest()


def run_sequential(cases, steps: int, chunk: int = 4096, ramp_steps: float = 3000, info: dict = None):
    """The same tunnels one after another as plain single-tunnel solvers (info as run_batched)."""
    elapsed = 0.0
    out = []
    div, sums = [], []
    for c in cases:
        s = Solver.create(dict(c["config"]))
        s.set_sdf(case_sdf(c))
        s.init_field()
        s.set_inlet_ramp("smoothstep", c["config"]["uIn"], ramp_steps)
        st, vs = [], []
        done = 0
        while done < steps:
            k = min(chunk, steps - done)
            elapsed += s.time_steps(k) / 1e3
            a, v = s.read_forces_raw()
            st.append(a)
            vs.append(v)
            done += k
        out.append((np.concatenate(st), np.concatenate(vs)))
        div.append(s.diverged_at())
        sums.append(_digest(s))
        s.destroy()
    if info is not None:
        info["diverged_at"] = div
        info["checksums"] = sums
    return elapsed, out


def compare_runs(cases, batched, sequential, window):
    """Batched against sequential runs of the same tunnels, divergence included. For every tunnel:
    the first non-finite force sample (the reference's blow-up rule, src/tunnel3d.ts:97-99) must fall
    on the same step in both runs and the force traces must agree bit for bit up to it (NaN payloads
    aside); mean C_D / C_L over the last `window` steps are compared on the tunnels that stayed
    finite. Returns a summary with the worst relative difference over those, the number of bit-equal
    traces and the diverged tunnels."""
    worst, equal, diverged, mismatch = 0.0, 0, [], []
    for i, (c, (sb, vb), (ss, vs)) in enumerate(zip(cases, batched, sequential)):
        if not np.array_equal(sb, ss):
            mismatch.append(dict(tunnel=i, what="sample steps differ"))
            continue
        bad_b = ~np.isfinite(vb).all(axis=1)
        bad_s = ~np.isfinite(vs).all(axis=1)
        first_b = int(sb[bad_b][0]) if bad_b.any() else 0
        first_s = int(ss[bad_s][0]) if bad_s.any() else 0
        upto = len(sb) if not first_b else int(np.argmax(bad_b))
        same = first_b == first_s and np.array_equal(vb[:upto], vs[:upto])
        equal += int(same and np.array_equal(vb, vs, equal_nan=True))
        if not same:
            mismatch.append(dict(tunnel=i, first_bad_batched=first_b, first_bad_sequential=first_s,
                                 max_abs_diff_before=float(np.abs(vb[:upto] - vs[:upto]).max()) if upto else 0.0))
        if first_b or first_s:
            cfg = c["config"]
            diverged.append(dict(tunnel=i, nx=cfg["width"], alpha_deg=round(c["alpha_deg"], 3), re=round(c["re"], 1),
                                 tau=cfg["tau"], smagorinsky=cfg["smagorinsky"], first_bad_step=first_b or first_s))
            continue
        cd_b, cl_b = coefficients(sb, vb, c["chord"], window)
        cd_s, cl_s = coefficients(ss, vs, c["chord"], window)
        worst = max(worst, abs(cd_b - cd_s) / max(abs(cd_s), 1e-12), abs(cl_b - cl_s) / max(abs(cl_s), 1e-3))
    return dict(tunnels=len(cases), finite_tunnels=len(cases) - len(diverged), bit_equal_traces=equal,
                mean_cd_cl_max_rel_diff_finite=worst, diverged=diverged, mismatches=mismatch)


def mlups(cases, steps, seconds):
    cells = sum(c["config"]["width"] * c["config"]["height"] for c in cases)
    return cells * steps / seconds / 1e6


if __name__ == "__main__":
    cs = sweep_cases()
    tb, rb = run_batched(cs, 2000)
    ts, rs = run_sequential(cs, 2000)
    print(f"batched {mlups(cs, 2000, tb):.0f} MLUPS, sequential {mlups(cs, 2000, ts):.0f} MLUPS, "
          f"ratio {ts / tb:.2f}; {compare_runs(cs, rb, rs, 1000)}")
