"""Writes tests/golden/adjoint_naca_384x128_fp64.json: the FP64 CPU reference (tests/adjoint_ref.py,
reverse mode through torch in float64) of the 2D shape-gradient case - D2Q9 384 x 128, NACA0012 of
chord 64 at 5 degrees with leading edge (96, 63.5), Re 400 (tau 0.548), Zou-He inlet u = 0.1 ramped
linearly over 3000 steps, Zou-He outlet, slip walls, 15% sponge, absorbing layer 0.02; J = mean C_D
over steps 4001..6000; parameters (alpha, a_1..a_8), bump width chord / 16. Takes about 15 minutes on
8 cores. Run from the repository root: python tests/golden/make_adjoint_naca_fp64.py"""
import json
import os
import sys
import time

import numpy as np
import torch

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
from cells_amd import shapes  # noqa: E402
from cells_amd.lbm import X_OPEN, Y_SLIP  # noqa: E402
from tests.adjoint_ref import reference_gradient  # noqa: E402

W, H, chord, u, T, window = 384, 128, 64.0, 0.1, 6000, 2000
tau = 3 * u * chord / 400 + 0.5
cfg = dict(width=W, height=H, left=X_OPEN, right=X_OPEN, bottom=Y_SLIP, top=Y_SLIP, tau=tau, uIn=u,
           spongeFraction=0.15, spongeTau=1, absorb=0.02, forceEvery=1)
theta = np.array([5.0] + [0.0] * 8)
sdf, dsdf = shapes.naca_bumps(W, H, 96.0, 63.5, chord, theta[0], theta[1:])
torch.set_num_threads(os.cpu_count() or 1)
t0 = time.time()
J, g = reference_gradient(cfg, sdf, T, window, u, 3000, chord, segment=20, ramp_kind="linear")
out = dict(case="naca_384x128", J=J, grad_theta=(dsdf @ g).tolist(), theta=theta.tolist(),
           seconds=time.time() - t0, nonzero_sdf_gradient=int(np.count_nonzero(g)))
path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "adjoint_naca_384x128_fp64.json")
with open(path, "w") as f:
    json.dump(out, f, indent=1)
print(json.dumps(out))
