"""Writes tests/golden/c1_sphere_192x96x96_fp64_steady.npy: the FP64 force trace (F_x, F_y, F_z per
step) over steps 19001..20000 of the same D3Q19 192 x 96 x 96 sphere tunnel as
make_c1_fp64_forces.py (oracle in double precision, inlet ramped by smoothstep over 3000 steps), by
which time C_D is steady (its spread over those 1000 steps is ~2e-8). The GPU precision study compares
the steady C_D of its FP32 / FP16 / BF16 storage runs with it. Run from the repository root after
building the oracle (about 30 minutes on 6 cores): python tests/golden/make_c1_fp64_steady.py"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import bench  # noqa: E402  (tunnel_config / tunnel_sdf define the workload)
from oracle import oracle  # noqa: E402

W, H, D, STEPS, KEEP = 192, 96, 96, 20000, 1000
s = oracle.OracleSolver3D(bench.tunnel_config(W, H, D, "fp32"), 64)
s.set_sdf(bench.tunnel_sdf(W, H, D))
s.init_field()
s.set_inlet_ramp("smoothstep", 0.1, 3000)
t0 = time.time()
steps, vals = [], []
for k in range(0, STEPS, 500):
    s.steps(500)
    st, v = s.read_forces_raw()
    if k + 500 > STEPS - KEEP:
        steps.append(st)
        vals.append(v)
    print(f"{k + 500} steps, {time.time() - t0:.0f} s", flush=True)
steps = np.concatenate(steps)
vals = np.concatenate(vals).reshape(-1, 3)
assert np.array_equal(steps, np.arange(STEPS - KEEP + 1, STEPS + 1)), "one sample per step expected"
np.save(os.path.join(ROOT, "tests", "golden", "c1_sphere_192x96x96_fp64_steady.npy"), vals)
print("done", vals[-1], time.time() - t0)
