"""Validation suite of the reference (src/cases.ts:61-427, src/cases3d.ts:11-81) run on this solver.

Every case is restated from the reference with its parameters, inlet ramp, convergence loop, metric
definitions and acceptance ranges unchanged (the reference fixes its ranges before a case runs,
AGENTS.md:90); the solver underneath is the gfx950 library through cells_amd.lbm. A case returns
{"name", "metrics": [{"name", "value", "ref", "accept", "pass"}], "pass", "steps", "cells",
"seconds", "notes"}.

Run from the repository root: python -m tests.validation.validate [case ...]   (default: every case)
"""
from __future__ import annotations

import json
import math
import sys
import time

import numpy as np

from cells_amd import geometry as g
from cells_amd.analysis import mean, period_from_crossings, primary_vortex, stream_function
from cells_amd.lbm import X_OPEN, X_PERIODIC, X_WALL, Y_MOVING, Y_NOSLIP, Y_PERIODIC, Y_SLIP, Solver, Solver3D


def metric(name, value, ref, accept):
    """metric (src/cases.ts:31-33)."""
    value = float(value)
    return dict(name=name, value=value, ref=list(ref), accept=list(accept),
                pass_=bool(accept[0] <= value <= accept[1]))


def around(center, rel):
    return [center * (1 - rel), center * (1 + rel)]


def run_ramped(s: Solver, steps: int, u_target: float, ramp_steps: float, forces: list, chunk: int = 400):
    """runRamped (src/cases.ts:41-59): smoothstep inlet ramp updated per chunk, forces collected."""
    end = s.step + steps
    while s.step < end:
        r = min(1.0, s.step / ramp_steps)
        u_in = u_target * r * r * (3 - 2 * r)
        if u_in != s.cfg["u_in"]:
            s.update(uIn=u_in)
        s.run(min(chunk, end - s.step), chunk)
        forces.extend(s.read_forces())
    return forces


def poiseuille():
    """src/cases.ts:61-92."""
    H, u_max = 32, 0.05
    metrics, steps = [], 0
    for tau in (0.6, 1.0, 1.5):
        nu = (tau - 0.5) / 3
        fx = 8 * nu * u_max / (H * H)
        s = Solver.create(dict(width=4, height=H, left=X_PERIODIC, right=X_PERIODIC, bottom=Y_NOSLIP, top=Y_NOSLIP,
                               tau=tau, bodyForce=(fx, 0.0), forceEvery=0))
        s.init_field()
        n = math.ceil(10 * H * H / nu)
        s.run(n, 2000)
        steps += n
        m = s.read_macro()
        y = np.arange(H)
        e = fx / (2 * nu) * (y + 0.5) * (H - y - 0.5)
        ux = m[(y * 4 + 1) * 4 + 1]
        err = math.sqrt(float(((ux - e) ** 2).sum() / (e * e).sum()))
        metrics.append(metric(f"L2 error, tau={tau}", err, [0, 0], [0, 2e-4]))
        s.destroy()
    return dict(name="Plane Poiseuille", metrics=metrics, steps=steps, cells=4 * H, notes=[])


def periodic_seam():
    """src/cases.ts:94-126."""
    W, H, D, tau = 128, 64, 16, 0.8
    drag, steps = [], 0
    for cx in (W / 2 + D / 2 - 0.5, D / 2 - 0.5):
        s = Solver.create(dict(width=W, height=H, left=X_PERIODIC, right=X_PERIODIC, bottom=Y_NOSLIP, top=Y_NOSLIP,
                               tau=tau, bodyForce=(2e-6, 0.0), forceEvery=100))
        sdf = g.empty_sdf(W, H)
        g.add_circle(sdf, W, H, cx, H / 2 - 0.5, D / 2)
        g.add_circle(sdf, W, H, cx + W, H / 2 - 0.5, D / 2)
        s.set_sdf(sdf)
        s.init_field()
        forces = []
        for _ in range(20):
            s.run(1000, 1000)
            forces.extend(s.read_forces())
        steps += s.step
        drag.append(mean(f.fx for f in forces[-5:]))
        s.destroy()
    return dict(name="Periodic seam invariance",
                metrics=[metric("drag at seam / drag in middle", drag[1] / drag[0], [1, 1], [0.999, 1.001])],
                steps=steps, cells=W * H, notes=[f"mid-domain drag {drag[0]:.7e}, seam drag {drag[1]:.7e}"])


def taylor_green():
    """src/cases.ts:128-192."""
    tau = 0.8
    nu = (tau - 0.5) / 3
    errors, nu_measured, steps = [], [], 0
    for n in (32, 64, 128):
        u0 = 0.04 * 32 / n
        k = 2 * math.pi / n
        s = Solver.create(dict(width=n, height=n, left=X_PERIODIC, right=X_PERIODIC, bottom=Y_PERIODIC,
                               top=Y_PERIODIC, tau=tau, forceEvery=0))
        ys, xs = np.mgrid[0:n, 0:n].astype(np.float64)

        def field(t):
            a = u0 * math.exp(-2 * nu * k * k * t)
            m = np.zeros((n, n, 4), dtype=np.float32)
            m[..., 0] = 1 - 0.75 * a * a * (np.cos(2 * k * xs) + np.cos(2 * k * ys))
            m[..., 1] = -a * np.cos(k * xs) * np.sin(k * ys)
            m[..., 2] = a * np.sin(k * xs) * np.cos(k * ys)
            return m.reshape(-1)

        s.init_field(field(0))
        t = round(math.log(2) / (2 * nu * k * k))

        def energy(m):
            m = m.astype(np.float64).reshape(-1, 4)
            return float((m[:, 1] ** 2 + m[:, 2] ** 2).sum())

        s.run(round(t / 2), 1000)
        mid = s.read_macro()
        s.run(t - round(t / 2), 1000)
        end = s.read_macro()
        steps += t
        exact = field(t).astype(np.float64).reshape(-1, 4)
        e = end.astype(np.float64).reshape(-1, 4)
        errors.append(math.sqrt(float(((e[:, 1:3] - exact[:, 1:3]) ** 2).sum() / (exact[:, 1:3] ** 2).sum())))
        dt = t - round(t / 2)
        nu_measured.append(math.log(energy(mid) / energy(end)) / (4 * k * k * dt))
        s.destroy()
    order = math.log2(errors[0] / errors[2]) / 2
    return dict(name="Taylor-Green vortex",
                metrics=[metric("nu measured / nu (N=128)", nu_measured[2] / nu, [1, 1], [0.99, 1.01]),
                         metric("velocity L2 error (N=128)", errors[2], [0, 0], [0, 0.01]),
                         metric("convergence order 32 to 128", order, [2, 2], [1.8, math.inf])],
                steps=steps, cells=128 * 128,
                notes=["L2 errors for N=32/64/128: " + ", ".join(f"{v:.2e}" for v in errors)])


def _schafer_turek(D, re, force_every, u_max):
    """schaferTurek (src/cases.ts:195-214)."""
    W, H = 22 * D, round(4.1 * D)
    u_mean = 2 / 3 * u_max
    nu = u_mean * D / re
    cx, cy = 2 * D, 2 * D - 0.5
    s = Solver.create(dict(width=W, height=H, left=X_OPEN, right=X_OPEN, bottom=Y_NOSLIP, top=Y_NOSLIP,
                           tau=3 * nu + 0.5, parabolic=True, uIn=0.0, forceEvery=force_every, spongeFraction=0.1,
                           spongeTau=1, absorb=0.1))
    sdf = g.empty_sdf(W, H)
    g.add_circle(sdf, W, H, cx, cy, D / 2)
    s.set_sdf(sdf)
    s.init_field()
    return s, W, H, cx, cy, u_mean, (lambda f: 2 * f / (u_mean * u_mean * D))


def schafer_turek_1():
    """src/cases.ts:216-269."""
    D = 40
    s, W, H, cx, cy, u_mean, coef = _schafer_turek(D, 20, 50, 0.05)
    forces, prev = [], math.inf
    for k in range(20):
        run_ramped(s, 10000, 0.05, 5000, forces)
        cd = coef(mean(f.fx for f in forces[-20:]))
        if k >= 4 and abs(cd - prev) < 1e-4:
            break
        prev = cd
    last = forces[-20:]
    cd = coef(mean(f.fx for f in last))
    cl = coef(mean(f.fy for f in last))
    m = s.read_macro().reshape(-1, 4)
    y0 = math.floor(cy)
    w = cy - y0

    def rho_at(x):
        return (1 - w) * m[y0 * W + x, 0] + w * m[(y0 + 1) * W + x, 0]

    def ux_at(x):
        return (1 - w) * m[y0 * W + x, 1] + w * m[(y0 + 1) * W + x, 1]

    xf, xb = cx - D / 2, cx + D / 2
    a, b = math.floor(xf), math.ceil(xb)
    p_front = rho_at(a) + (rho_at(a) - rho_at(a - 1)) * (xf - a)
    p_back = rho_at(b) + (rho_at(b) - rho_at(b + 1)) * (b - xb)
    dp = (p_front - p_back) / 3 / (u_mean * u_mean)
    la = math.nan
    for x in range(b, W - 1):
        u0, u1 = ux_at(x), ux_at(x + 1)
        if u0 < 0 <= u1:
            la = (x + -u0 / (u1 - u0) - xb) / D
            break
    steps = s.step
    s.destroy()
    return dict(name="Schäfer-Turek 2D-1 (Re 20)",
                metrics=[metric("c_D", cd, [5.57, 5.59], around(5.58, 0.02)),
                         metric("c_L", cl, [0.0104, 0.011], around(0.0107, 0.1)),
                         metric("L_a / D", la, [0.842, 0.852], around(0.847, 0.03)),
                         metric("dP / (rho U_mean^2)", dp, [0.1172 / 0.04, 0.1176 / 0.04], around(0.1174 / 0.04, 0.02))],
                steps=steps, cells=W * H, notes=[f"D = {D} cells, grid {W}x{H}"])


def schafer_turek_2():
    """src/cases.ts:271-306."""
    D = 40
    s, W, H, cx, cy, u_mean, coef = _schafer_turek(D, 100, 5, 0.05)
    period = D / (0.3 * u_mean)
    forces, window, prev_amp = [], [], 0.0
    for k in range(40):
        run_ramped(s, round(5 * period), 0.05, 5000, forces)
        window = [f for f in forces if f.step > s.step - 5 * period]
        cls = [coef(f.fy) for f in window]
        amp = max(cls) - min(cls)
        if k >= 6 and amp > 0.5 and abs(amp - prev_amp) < 1e-3 * amp:
            break
        prev_amp = amp
        del forces[: max(0, len(forces) - 20000)]
    t = [f.step for f in window]
    cl = [coef(f.fy) for f in window]
    cd = [coef(f.fx) for f in window]
    p = period_from_crossings(t, cl)
    st = D / (p[0] * u_mean) if p else math.nan
    steps = s.step
    s.destroy()
    return dict(name="Schäfer-Turek 2D-2 (Re 100)",
                metrics=[metric("c_D,max", max(cd), [3.22, 3.24], around(3.23, 0.02)),
                         metric("c_L,max", max(cl), [0.99, 1.01], around(1.0, 0.05)),
                         metric("St", st, [0.295, 0.305], around(0.3, 0.02))],
                steps=steps, cells=W * H, notes=[f"D = {D} cells, grid {W}x{H}, forces every 5 steps"])


def unconfined_cylinder():
    """src/cases.ts:308-355."""
    D, lx, ly, xc, u = 32, 40, 48, 16, 0.05
    W, H = lx * D, ly * D
    nu = u * D / 100
    s = Solver.create(dict(width=W, height=H, left=X_OPEN, right=X_OPEN, bottom=Y_SLIP, top=Y_SLIP,
                           tau=3 * nu + 0.5, forceEvery=5, spongeFraction=0.15, spongeTau=1, absorb=0.02))
    sdf = g.empty_sdf(W, H)
    g.add_circle(sdf, W, H, xc * D, H / 2 + 0.3, D / 2)
    s.set_sdf(sdf)
    s.init_field()

    def coef(f):
        return 2 * f / (u * u * D)

    period = D / (0.165 * u)
    forces, window, prev_amp = [], [], 0.0
    for k in range(60):
        run_ramped(s, round(5 * period), u, 5000, forces)
        window = [f for f in forces if f.step > s.step - 5 * period]
        cls = [coef(f.fy) for f in window]
        amp = max(cls) - min(cls)
        if k >= 6 and amp > 0.3 and abs(amp - prev_amp) < 2e-3 * amp:
            break
        prev_amp = amp
        del forces[: max(0, len(forces) - 20000)]
    t = [f.step for f in window]
    p = period_from_crossings(t, [f.fy for f in window])
    st = D / (p[0] * u) if p else math.nan
    n = math.floor((t[-1] - t[0]) / p[0]) if p else 0
    span = [f for f in window if f.step > t[-1] - n * (p[0] if p else 0)]
    cd = coef(mean(f.fx for f in span))
    steps = s.step
    s.destroy()
    return dict(name="Unconfined cylinder (Re 100)",
                metrics=[metric("St", st, [0.1643, 0.1670], around(0.165, 0.04)),
                         metric("mean C_D", cd, [1.33, 1.36], around(1.345, 0.05))],
                steps=steps, cells=W * H, notes=[f"D = {D} cells, domain {lx}D x {ly}D, slip side walls"])


def cavity():
    """src/cases.ts:357-390."""
    N, u = 128, 0.1
    nu = u * N / 1000
    s = Solver.create(dict(width=N, height=N, left=X_WALL, right=X_WALL, bottom=Y_NOSLIP, top=Y_MOVING,
                           tau=3 * nu + 0.5, uLid=u, forceEvery=0))
    s.init_field()
    v, prev = (0.0, 0.0, 0.0), (-1.0, -1.0)
    for k in range(40):
        s.run(20000, 2000)
        m = s.read_macro().reshape(-1, 4)
        psi = stream_function(m[:, 1], m[:, 2], N, N)
        v = primary_vortex(psi, N, N)
        if math.hypot(v[0] - prev[0], v[1] - prev[1]) < 0.01 and k >= 5:
            break
        prev = v
    steps = s.step
    s.destroy()
    return dict(name="Ghia lid-driven cavity (Re 1000)",
                metrics=[metric("vortex x / L", (v[0] + 0.5) / N, [0.5313, 0.5313], [0.5213, 0.5413]),
                         metric("vortex y / L", (v[1] + 0.5) / N, [0.5625, 0.5625], [0.5525, 0.5725])],
                steps=steps, cells=N * N, notes=[f"N = {N}, lid u = {u}"])


def naca():
    """src/cases.ts:392-425."""
    c, u = 100, 0.1
    W, H = 16 * c, 8 * c
    nu = u * c / 500
    s = Solver.create(dict(width=W, height=H, left=X_OPEN, right=X_OPEN, bottom=Y_SLIP, top=Y_SLIP,
                           tau=3 * nu + 0.5, forceEvery=50, spongeFraction=0.15, spongeTau=1, absorb=0.1,
                           inletLayerFraction=0.08))
    sdf = g.empty_sdf(W, H)
    g.add_polygon(sdf, W, H, g.naca_polygon(4 * c, H / 2 - 0.5, c, 0.12, 0.0))
    s.set_sdf(sdf)
    s.init_field()
    forces, prev, cd = [], math.inf, math.nan
    for k in range(30):
        run_ramped(s, 10000, u, 5000, forces)
        cd = 2 * mean(f.fx for f in forces[-20:]) / (u * u * c)
        if k >= 5 and abs(cd - prev) < 2e-4:
            break
        prev = cd
    steps = s.step
    s.destroy()
    return dict(name="NACA0012 (Re 500, alpha 0)",
                metrics=[metric("C_D", cd, [0.1762, 0.178], around(0.1762, 0.05))],
                steps=steps, cells=W * H, notes=[f"chord {c} cells, domain 16c x 8c"])


def sphere_fp16():
    """src/cases3d.ts:11-81: sphere C_D at Re 100 in FP32 and FP16 storage, 192 x 96 x 96."""
    W, H, D = 192, 96, 96
    d, u = 16, 0.1
    from cells_amd.analysis import derive_tau

    tau = derive_tau(100, u, d).tau
    sdf = g.empty_sdf3(W, H, D)
    g.add_sphere(sdf, W, H, D, W / 4, H / 2, D / 2, d / 2)
    scale = 2 / (u * u * g.reference_area("sphere", d, D))
    cd, steps, notes = {}, 0, []
    for precision in ("fp32", "fp16"):
        s = Solver3D.create(dict(width=W, height=H, depth=D, precision=precision, tau=tau, spongeFraction=0.15,
                                 spongeTau=1, absorb=0.02, forceEvery=4))
        s.set_sdf(sdf)
        s.init_field()
        prev, converged = math.nan, False
        for k in range(1, 16):
            block, end = [], k * 2000
            while s.step < end:
                r = min(1.0, s.step / 3000)
                s.set_inlet(u * r * r * (3 - 2 * r))
                s.run(min(400, end - s.step), 400)
                block.extend(f.fx for f in s.read_forces())
            c = mean(block) * scale
            cd[precision] = c
            if k >= 5 and abs(c - prev) / abs(prev) < 1e-4:
                converged = True
                break
            prev = c
        if not converged:
            notes.append(f"{precision} reached 15 blocks without converging")
        steps += s.step
        s.destroy()
    notes.append(f"C_D FP32 {cd['fp32']:.5f}, FP16 {cd['fp16']:.5f}")
    return dict(name="Sphere C_D, FP16 vs FP32 (Re 100)",
                metrics=[metric("C_D relative difference FP16 vs FP32", abs(cd["fp16"] - cd["fp32"]) / cd["fp32"],
                                [0, 0], [0, 0.01])],
                steps=steps, cells=W * H * D, notes=notes)


CASES = {
    "poiseuille": poiseuille,
    "periodicSeam": periodic_seam,
    "taylorGreen": taylor_green,
    "schaferTurek1": schafer_turek_1,
    "schaferTurek2": schafer_turek_2,
    "unconfinedCylinder": unconfined_cylinder,
    "cavity": cavity,
    "naca": naca,
    "sphereFp16": sphere_fp16,
}


def run_case(name: str) -> dict:
    t0 = time.perf_counter()
    r = CASES[name]()
    r["seconds"] = time.perf_counter() - t0
    r["pass"] = all(m["pass_"] for m in r["metrics"]) and bool(r["metrics"])
    return r


def main(argv):
    names = argv or list(CASES)
    results = []
    for n in names:
        r = run_case(n)
        results.append(r)
        print(json.dumps(r), flush=True)
    return 0 if all(r["pass"] for r in results) else 1


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
