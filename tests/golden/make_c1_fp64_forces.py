"""Writes tests/golden/c1_sphere_192x96x96_fp64_forces.npy: the FP64 force trace (F_x, F_y, F_z per
step, steps 1..4000) of the D3Q19 192 x 96 x 96 sphere tunnel (sphere of diameter 16 at (48, 48, 48),
tau from Re 200 at u = 0.1 over 16 cells, Lambda = min(3/16, 50 (tau - 1/2)^2), equilibrium inlet
u = 0.1 ramped by smoothstep over 3000 steps, pressure outlet, 15% sponge, absorbing layer 0.02,
slip side faces, initial field from lbm_init_field), run by the oracle in double precision. The GPU
precision study compares its FP32 / FP16 / BF16 storage runs of the same tunnel with it. Run from
the repository root after building the oracle: python tests/golden/make_c1_fp64_forces.py"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import bench  # noqa: E402  (tunnel_config / tunnel_sdf define the workload)
from oracle import oracle  # noqa: E402

W, H, D, STEPS = 192, 96, 96, 4000
s = oracle.OracleSolver3D(bench.tunnel_config(W, H, D, "fp32"), 64)
s.set_sdf(bench.tunnel_sdf(W, H, D))
s.init_field()
s.set_inlet_ramp("smoothstep", 0.1, 3000)
t0 = time.time()
steps, vals = [], []
for k in range(0, STEPS, 500):
    s.steps(500)
    st, v = s.read_forces_raw()
    steps.append(st)
    vals.append(v)
    print(f"{k + 500} steps, {time.time() - t0:.0f} s", flush=True)
steps = np.concatenate(steps)
vals = np.concatenate(vals).reshape(-1, 3)
assert np.array_equal(steps, np.arange(1, STEPS + 1)), "one sample per step expected"
np.save(os.path.join(ROOT, "tests", "golden", "c1_sphere_192x96x96_fp64_forces.npy"), vals)
print("done", vals[-1], time.time() - t0)
