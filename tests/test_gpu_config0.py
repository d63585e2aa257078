"""Config 0 end to end: the D2Q9 512x128 cylinder tunnel for 20 000 steps on the GPU against the
FP32 and FP64 oracles (the reference method: drag/lift traces and the Strouhal number).

Setup (with the reference's tunnel defaults src/app.ts:176-190): TRT Lambda = 3/16,
FP32 f - w storage, circle of diameter 16 at (128, 64.5) with Bouzidi walls, Zou-He inlet u = 0.1
ramped linearly over 3000 steps, Zou-He pressure outlet, slip top/bottom, sponge over the last 15%
(spongeTau 1, absorbing layer 0.02), Re = 100 -> nu = 0.016, tau = 0.548, no Smagorinsky.
C = 2 F / (u^2 D); St = D / (period u) from upward crossings of C_L's mean over the second half.

Tolerances: the flow sheds vortices, so pointwise traces of two FP32 computations separate slowly;
the per-quantity gates are on the statistics: mean C_D within 0.5%, C_L amplitude within 2%,
St within 1% (GPU vs FP32 oracle); the same against FP64 are printed. This geometry is confined
(12.5% blockage, 8 D of upstream run), so C_D and St sit a few percent above the unconfined
reference ranges (src/cases.ts:346-347); that is reported, not gated.
"""
from __future__ import annotations

import numpy as np
import pytest

from cells_amd import geometry as g
from cells_amd.analysis import period_from_crossings
from cells_amd.lbm import X_OPEN, Y_SLIP, Solver
from oracle.oracle import OracleSolver

pytestmark = pytest.mark.gpu

W, H, D, U = 512, 128, 16, 0.1
STEPS = 20000


def config0():
    return dict(width=W, height=H, left=X_OPEN, right=X_OPEN, bottom=Y_SLIP, top=Y_SLIP, tau=3 * 0.016 + 0.5,
                spongeFraction=0.15, spongeTau=1, absorb=0.02, forceEvery=1)


def _run(sim, chunk):
    sdf = g.empty_sdf(W, H)
    g.add_circle(sdf, W, H, 128, 64.5, D / 2)
    sim.set_sdf(sdf)
    sim.init_field()
    sim.set_inlet_ramp("linear", U, 3000)
    steps, vals = [], []
    done = 0
    while done < STEPS:
        k = min(chunk, STEPS - done)
        sim.run(k) if hasattr(sim, "sync") else sim.steps(k)
        s, v = sim.read_forces_raw()
        steps.append(s)
        vals.append(v)
        done += k
    return np.concatenate(steps), np.concatenate(vals)


def _stats(steps, vals):
    c = 2 * vals / (U * U * D)
    half = steps > STEPS // 2
    cd = c[half, 0].mean()
    cl = c[half, 1]
    p = period_from_crossings(steps[half], cl)
    st = D / (p[0] * U) if p else float("nan")
    return cd, cl.max() - cl.min(), st


@pytest.mark.timeout(400)
def test_config0_cd_cl_strouhal_match_oracle():
    sg, vg = _run(Solver.create(config0()), 4000)
    so, vo = _run(OracleSolver(config0(), 32), 4000)
    s6, v6 = _run(OracleSolver(config0(), 64), 4000)
    assert len(sg) == STEPS and (sg == so).all() and (sg == s6).all()
    cd_g, amp_g, st_g = _stats(sg, vg)
    cd_o, amp_o, st_o = _stats(so, vo)
    cd_6, amp_6, st_6 = _stats(s6, v6)
    print(f"GPU   C_D {cd_g:.5f}  C_L amp {amp_g:.5f}  St {st_g:.5f}")
    print(f"o32   C_D {cd_o:.5f}  C_L amp {amp_o:.5f}  St {st_o:.5f}")
    print(f"o64   C_D {cd_6:.5f}  C_L amp {amp_6:.5f}  St {st_6:.5f}")
    print(f"GPU vs FP64: C_D {abs(cd_g - cd_6) / cd_6:.2e}, St {abs(st_g - st_6) / st_6:.2e}")
    print(f"unconfined reference ranges: St [0.1584, 0.1716], C_D [1.278, 1.412] (this tunnel is confined)")
    # the early, still-laminar part of the trace agrees closely pointwise
    early = (sg > 500) & (sg <= 4000)
    rel = np.abs(vg[early] - vo[early]).max() / np.abs(vo[early]).max()
    print(f"max rel force difference, steps 500-4000: {rel:.2e}")
    assert rel < 1e-3
    assert abs(cd_g - cd_o) / cd_o < 5e-3
    assert abs(amp_g - amp_o) / amp_o < 2e-2
    assert abs(st_g - st_o) / st_o < 1e-2
    assert 0.15 < st_g < 0.2 and 1.2 < cd_g < 1.7
