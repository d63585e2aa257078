"""Writes tests/golden/adjoint_fd_fp64.json: central finite differences of the two shape objectives
computed with the FP64 oracle (the C++ restatement in double precision, forward only), the FP64
reference the GPU's discrete adjoint is checked against besides the reverse-mode reference of
tests/adjoint_ref.py.

Cases (each objective is run exactly as lbm_adjoint_forward runs it: the configured inlet ramp, the
initial field at rest, one force sample per step, J = scale * mean over the samples of steps
window_start + 1 .. steps of direction . F):
  naca_384x128   D2Q9 384 x 128, NACA0012 of chord 64 at 5 degrees with leading edge (96, 63.5), Re 400
                 (tau 0.548), Zou-He inlet u = 0.1 ramped linearly over 3000 steps, Zou-He outlet,
                 slip walls, 15% sponge, absorbing layer 0.02; J = mean C_D over steps 4001..6000;
                 parameters (alpha, a_1..a_8) of cells_amd.shapes.naca_bumps.
  sphere_48x24x24  D3Q19 48 x 24 x 24, sphere of radius 5 at (12, 11.5, 11.5) with 16 real spherical
                 harmonics (l <= 3) on its radius, tau 0.56, equilibrium inlet u = 0.05 ramped by
                 smoothstep over 100 steps, 15% sponge, absorbing layer 0.02; J = mean C_D over steps
                 201..300 (area pi r^2).
For every parameter and step h in (1e-3, 1e-4): (J(theta + h e_p) - J(theta - h e_p)) / 2h, and the
number of nodes whose solid / fluid flag differs between theta +- h e_p and theta (a flip makes the
objective jump, which a difference quotient would mistake for a slope). About 2 minutes on 8 cores.
Run from the repository root: python tests/golden/make_adjoint_fd_fp64.py
"""
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from cells_amd import shapes  # noqa: E402
from cells_amd.lbm import X_OPEN, Y_SLIP  # noqa: E402
from oracle.oracle import OracleSolver, OracleSolver3D  # noqa: E402

STEPS_H = (1e-3, 1e-4)


def objective(make, cfg, sdf, T, window, direction, scale, ramp):
    o = make(dict(cfg), 64)
    o.set_sdf(sdf)
    kind, u, rs = ramp
    o.set_inlet_ramp(kind, u, rs)
    if isinstance(o, OracleSolver3D):
        o.set_inlet(0.0)  # the 3D initial field is the equilibrium at the first step's speed: rest
    o.init_field()
    o.steps(T)
    st, v = o.read_forces_raw()
    o.destroy()
    assert len(st) == T and st[0] == 1 and st[-1] == T
    sel = st > T - window
    return float(scale * (v[sel] @ np.asarray(direction[: v.shape[1]], dtype=np.float64)).mean())


def solid_of(sdf, W):
    """Solid nodes as the flag kernels decide them: sdf < 0 except the open inlet / outlet planes."""
    s = np.asarray(sdf, dtype=np.float32).reshape(-1, W) < 0
    s[:, 0] = s[:, -1] = False
    return s.reshape(-1)


def fd_case(name, make, cfg, sdf_of, theta, T, window, direction, scale, ramp):
    t0 = time.time()
    sdf0, _ = sdf_of(theta)
    solid0 = solid_of(sdf0, cfg["width"])
    J0 = objective(make, cfg, sdf0, T, window, direction, scale, ramp)
    out = dict(case=name, J=J0, theta=list(map(float, theta)), T=T, window=window)
    for h in STEPS_H:
        fd, flips = [], []
        for p in range(len(theta)):
            e = np.zeros_like(theta)
            e[p] = h
            sp, sm = sdf_of(theta + e)[0], sdf_of(theta - e)[0]
            flips.append(int(np.count_nonzero(solid_of(sp, cfg["width"]) != solid0) +
                             np.count_nonzero(solid_of(sm, cfg["width"]) != solid0)))
            jp = objective(make, cfg, sp, T, window, direction, scale, ramp)
            jm = objective(make, cfg, sm, T, window, direction, scale, ramp)
            fd.append((jp - jm) / (2 * h))
        out[f"fd_h{h:g}"] = fd
        out[f"flips_h{h:g}"] = flips
    out["seconds"] = time.time() - t0
    return out


def main():
    cases = {}
    # 2D shape-gradient case
    W, H, chord, u = 384, 128, 64.0, 0.1
    tau = 3 * u * chord / 400 + 0.5
    cfg = dict(width=W, height=H, left=X_OPEN, right=X_OPEN, bottom=Y_SLIP, top=Y_SLIP, tau=tau, uIn=u,
               spongeFraction=0.15, spongeTau=1, absorb=0.02, forceEvery=1)
    cases["naca_384x128"] = fd_case(
        "naca_384x128", OracleSolver, cfg,
        lambda th: shapes.naca_bumps(W, H, 96.0, 63.5, chord, th[0], th[1:]),
        np.array([5.0] + [0.0] * 8), 6000, 2000, (1.0, 0.0), 2 / (u * u * chord), ("linear", u, 3000))
    # small 3D sphere case
    W3, H3, D3, u3, r = 48, 24, 24, 0.05, 5.0
    cfg3 = dict(width=W3, height=H3, depth=D3, precision="fp32", tau=0.56, uIn=u3, spongeFraction=0.15, spongeTau=1,
                absorb=0.02, forceEvery=1)
    cases["sphere_48x24x24"] = fd_case(
        "sphere_48x24x24", OracleSolver3D, cfg3,
        lambda b: shapes.sphere_sh(W3, H3, D3, (12.0, 11.5, 11.5), r, b),
        np.zeros(16), 300, 100, (1.0, 0.0, 0.0), 2 / (u3 * u3 * math.pi * r * r), ("smoothstep", u3, 100))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "adjoint_fd_fp64.json")
    with open(path, "w") as f:
        json.dump(cases, f, indent=1)
    print(json.dumps(cases))


if __name__ == "__main__":
    main()
