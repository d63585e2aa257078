"""Storage precision on the D3Q19 192 x 96 x 96 sphere tunnel: the C_D trace of FP32, FP16 x 2^15 and
BF16 storage runs against the FP64 run of the same tunnel at identical steps (the oracle in double
precision, committed by tests/golden/make_c1_fp64_forces.py and make_c1_fp64_steady.py), as window
means over steps 3001-4000 (the wake still developing after the inlet ramp) and 19001-20000 (steady
C_D). FP32 agrees to 1e-4; FP16 x 2^15 (11-bit mantissa of the scaled f - w) and BF16 (8-bit mantissa,
no scaling) are held to 1% (measured: FP16 0.67% and 0.82%, BF16 0.65% and 0.13%); the same FP16
format with stochastic rounding in the step (LBM_FP16_SR) to 0.5%: rounding to nearest repeats the
same rounding error at a node step after step once the flow is steady, and the errors add up to a
bias that stochastic rounding does not have.

The FP16 figure is the storage format's, not the kernels': on a smaller sphere tunnel (64 x 32 x 32,
5000 steps) run to a steady state, the GPU's FP16 C_D equals the FP16 oracle's (same format, same
rounding) to 5e-4 of its value, a fraction of the FP16-vs-FP64 gap there (0.11%)."""
import math
import os

import numpy as np
import pytest

import bench

pytestmark = pytest.mark.gpu


@pytest.mark.skipif(not (os.path.exists(bench.PRECISION_FIXTURE) and os.path.exists(bench.STEADY_FIXTURE)),
                    reason="FP64 force traces not generated")
def test_storage_precision_against_fp64():
    res = bench.precision_study(20000)
    for prec, r in res.items():
        print(prec, r)
    for window in ("transient_3001_4000", "steady_19001_20000"):
        assert res["fp32"][window]["rel_err_mean"] < 1e-4
        assert res["fp16"][window]["rel_err_mean"] < 1e-2
        assert res["fp16sr"][window]["rel_err_mean"] < 5e-3
        assert res["bf16"][window]["rel_err_mean"] < 1e-2


def test_fp16_steady_drag_matches_fp16_oracle():
    from cells_amd import geometry as g
    from cells_amd.lbm import Solver3D
    from oracle.oracle import OracleSolver3D, set_threads

    W, H, D, r, u, steps = 64, 32, 32, 3.0, 0.1, 5000
    cfg = dict(width=W, height=H, depth=D, precision="fp16", tau=0.524, uIn=0.0, spongeFraction=0.15, spongeTau=1,
               absorb=0.02, forceEvery=1, workgroupSize=256)
    sdf = g.empty_sdf3(W, H, D)
    g.add_sphere(sdf, W, H, D, W / 4, H / 2, D / 2, r)
    scale = 2 / (u * u * math.pi * r * r)
    set_threads(min(16, os.cpu_count() or 1))
    cd = {}
    for name, sim in (("gpu", Solver3D.create(dict(cfg))), ("o16", OracleSolver3D(dict(cfg), 32)),
                      ("o64", OracleSolver3D(dict(cfg, precision="fp32"), 64))):
        sim.set_sdf(sdf)
        sim.init_field()
        sim.set_inlet_ramp("smoothstep", u, 1500)
        vals = []
        for k in range(0, steps, 1000):
            if name == "gpu":
                sim.encode_steps(1000)
            else:
                sim.steps(1000)
            vals.append(sim.read_forces_raw()[1])
        sim.destroy()
        cd[name] = scale * np.concatenate(vals).reshape(-1, 3)[-500:, 0].mean()
    print({k: round(float(v), 7) for k, v in cd.items()})
    assert abs(cd["gpu"] / cd["o16"] - 1) < 5e-4
    assert abs(cd["o16"] / cd["o64"] - 1) < 1e-2
