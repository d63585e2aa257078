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


def test_batch_launch_shapes_give_identical_bits():
    """Every tile width, tunnel order and bulk-tile direction the batch autotuner picks from stores
    the same bits and the same force samples (cells are independent within a step; force samples are
    summed per wave of body nodes in node order, whatever the block size or the order blocks run in),
    also across the graph-captured 32-step loops (250 steps, chunks of 100)."""
    cases = sweep.sweep_cases(n=5, seed=7, sizes=(64, 96, 160))
    ref = None
    for tile, order in ((256, 0), (64, 0), (128, 1), (512, 1), (256, 2), (64, 3), (512, 2)):
        info = {}
        _, traces = sweep.run_batched(cases, 250, chunk=100, ramp_steps=100, launch=(tile, order), info=info)
        assert (info["tile"], info["order"]) == (tile, order)
        got = (info["checksums"], [v for _, v in traces])
        if ref is None:
            ref = got
            continue
        assert got[0] == ref[0], (tile, order)
        for a, b in zip(got[1], ref[1]):
            np.testing.assert_array_equal(a, b)


def test_sweep_64_tunnels_seed0_full_size():
    """The 64-tunnel seed-0 NACA0012 sweep at full size (7.43 M cells) for 2000 steps, before the
    earliest divergence of the marginally stable tunnels: the autotuned batch against the same tunnels
    run one by one (force traces and final populations bit for bit, no tunnel flagged as diverged),
    and the smallest and the largest tunnel's force samples against the FP32 oracle."""
    from oracle.oracle import OracleSolver

    cases = sweep.sweep_cases(64, seed=0)
    steps = 2000
    ib, iseq = {}, {}
    tb, rb = sweep.run_batched(cases, steps, chunk=1000, info=ib)
    ts, rs = sweep.run_sequential(cases, steps, chunk=1000, info=iseq)
    cmp = sweep.compare_runs(cases, rb, rs, 1000)
    print(f"batched {sweep.mlups(cases, steps, tb):.0f} MLUPS (tile {ib['tile']}, order {ib['order']}), sequential "
          f"{sweep.mlups(cases, steps, ts):.0f} MLUPS; {cmp['bit_equal_traces']}/64 traces bit-equal")
    assert not cmp["mismatches"], cmp["mismatches"]
    assert cmp["bit_equal_traces"] == 64 and cmp["finite_tunnels"] == 64
    assert ib["diverged_at"] == [0] * 64 and iseq["diverged_at"] == [0] * 64
    assert ib["checksums"] == iseq["checksums"]
    widths = [c["config"]["width"] for c in cases]
    for i in (widths.index(min(widths)), widths.index(max(widths))):
        c = cases[i]
        o = OracleSolver(dict(c["config"]), 32)
        o.set_sdf(sweep.case_sdf(c))
        o.init_field()
        o.set_inlet_ramp("smoothstep", c["config"]["uIn"], 3000)
        o.steps(steps)
        so, vo = o.read_forces_raw()
        sb, vb = rb[i]
        np.testing.assert_array_equal(so, sb)
        rel = np.abs(vb - vo).max() / np.abs(vo).max()
        print(f"tunnel {i} ({c['config']['width']}x{c['config']['height']}): force rel |gpu-o32| {rel:.2e}")
        assert rel < 1e-4
        o.destroy()


def test_divergence_is_recorded_per_tunnel():
    """hasDiverged (src/tunnel3d.ts:97-99) kept on the device: a tunnel driven unstable (tau 0.5005,
    u 0.3, no sponge) records the step of its first non-finite force sample, the same step as its
    force trace and as the oracle's; the stable tunnels of the same batch record none; resetForces
    clears the record, and the next sample of the blown-up flow sets it again."""
    from cells_amd.lbm import Batch, Solver
    from oracle.oracle import OracleSolver
    from tests.parity_cases import CASES2

    cfg, fn = CASES2["tunnel"]
    unstable = dict(cfg, tau=0.5005, uIn=0.3, absorb=0.0, spongeFraction=0.0)
    W, H = cfg["width"], cfg["height"]
    o = OracleSolver(dict(unstable), 32)
    o.set_sdf(fn(W, H))
    o.init_field()
    o.steps(300)
    so, vo = o.read_forces_raw()
    first_oracle = int(so[~np.isfinite(vo).all(1)][0])
    b = Batch.create([dict(cfg), dict(unstable), dict(cfg)])
    for t in b.tunnels:
        t.set_sdf(fn(W, H))
        t.init_field()
    b.reset_clock()
    b.run(300)
    d = b.diverged_at()
    s1, v1 = b.tunnels[1].read_forces_raw()
    first_trace = int(s1[~np.isfinite(v1).all(1)][0])
    print(f"diverged at {list(d)}; trace {first_trace}, oracle {first_oracle}")
    assert d[0] == 0 and d[2] == 0
    assert d[1] == first_trace
    assert abs(first_trace - first_oracle) <= 5  # an instability, not a bit-level quantity
    assert b.tunnels[1].has_diverged() and not b.tunnels[0].has_diverged()
    b.tunnels[1].reset_forces()
    assert b.tunnels[1].diverged_at() == 0
    b.run(1)
    assert b.tunnels[1].diverged_at() == 301
    b.destroy()
    s = Solver.create(dict(unstable))
    s.set_sdf(fn(W, H))
    s.init_field()
    s.run(300)
    assert s.diverged_at() == first_trace
    s.init_field()
    assert s.diverged_at() == 0
    s.destroy()
