"""Batched tunnels (one launch per step for all, hipGraph replay) against the same tunnels run alone:
per-tunnel force traces and mean C_D / C_L agree (the per-cell code, the force tiles and the FP64
summation order are the same, so agreement is expected to the last bit; the gate is 1e-6 relative)."""
from __future__ import annotations

import numpy as np
import pytest

from cells_amd import sweep

pytestmark = pytest.mark.gpu


def test_batch_matches_separate_runs():
    cases = sweep.sweep_cases(n=7, seed=3, sizes=(64, 96, 128, 160))
    steps = 1200  # not a multiple of the graph length: exercises graph replays and direct launches
    tb, rb = sweep.run_batched(cases, steps, chunk=500, ramp_steps=400)
    ts, rs = sweep.run_sequential(cases, steps, chunk=500, ramp_steps=400)
    exact = 0
    for c, (sb, vb), (ss, vs) in zip(cases, rb, rs):
        assert len(sb) == steps and (sb == ss).all()
        assert np.isfinite(vb).all()
        scale = np.abs(vs).max()
        rel = np.abs(vb - vs).max() / scale
        exact += int(rel == 0.0)
        cd_b, cl_b = sweep.coefficients(sb, vb, c["chord"], 400)
        cd_s, cl_s = sweep.coefficients(ss, vs, c["chord"], 400)
        print(f"{c['config']['width']}x{c['config']['height']} Re {c['re']:.0f} a {c['alpha_deg']:.1f}: "
              f"C_D {cd_b:.5f}/{cd_s:.5f} C_L {cl_b:.5f}/{cl_s:.5f} max rel {rel:.1e}")
        assert rel < 1e-6
        assert abs(cd_b - cd_s) <= 1e-6 * abs(cd_s) + 1e-12
    print(f"{exact}/{len(cases)} tunnels bit-identical; batched {sweep.mlups(cases, steps, tb):.0f} MLUPS, "
          f"sequential {sweep.mlups(cases, steps, ts):.0f} MLUPS")


def test_batch_populations_match_separate_run():
    cases = sweep.sweep_cases(n=3, seed=5, sizes=(64, 128))
    from cells_amd.lbm import Batch, Solver

    b = Batch.create([c["config"] for c in cases])
    for i, c in enumerate(cases):
        b.tunnel(i).set_sdf(sweep.case_sdf(c))
        b.tunnel(i).init_field()
    b.reset_clock()
    b.run(77)
    for i, c in enumerate(cases):
        s = Solver.create(dict(c["config"]))
        s.set_sdf(sweep.case_sdf(c))
        s.init_field()
        s.run(77)
        fb = b.tunnels[i].read_populations()
        fs = s.read_populations()
        assert np.abs(fb - fs).max() <= 1e-6 * max(1.0, np.abs(fs).max())
        assert b.tunnels[i].step == 77
    b.destroy()
