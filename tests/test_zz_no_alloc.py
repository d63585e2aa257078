"""lbm_enqueue_steps_noalloc (the allocation-free variant of lbm_enqueue_steps, whose force reduction
keeps its scratch in the library's static per-GPU array) against lbm_enqueue_steps on a twin
handle: populations and force samples bit for bit, on two calls in a row, the second on a smaller
tunnel (fewer body-link waves, so partials the first call left behind must not leak into its
samples); a tunnel whose reduction needs more partials than the array holds is refused."""
from __future__ import annotations

import numpy as np
import pytest

from cells_amd import geometry as g
from cells_amd.lbm import LbmError, Solver, Solver3D
from tests.parity_cases import CASES2, case3

pytestmark = pytest.mark.gpu


def _pair(make):
    a, b = make(), make()
    return a, b


def _same(a, b):
    np.testing.assert_array_equal(a.read_populations_raw(), b.read_populations_raw())
    sa, va = a.read_forces_raw()
    sb, vb = b.read_forces_raw()
    np.testing.assert_array_equal(sa, sb)
    np.testing.assert_array_equal(va, vb)
    return len(sa)


def test_noalloc_steps_match_enqueue_steps():
    # first call: a D3Q19 tunnel with a large sphere (many body-link waves), FP32
    cfg, _ = case3("fp32", W=48, H=24, D=24)
    W, H, D = cfg["width"], cfg["height"], cfg["depth"]
    sdf = g.empty_sdf3(W, H, D)
    g.add_sphere(sdf, W, H, D, 16, 11.5, 11.5, 7.3)

    def big():
        s = Solver3D.create(dict(cfg))
        s.set_sdf(sdf)
        s.init_field()
        return s

    a, b = _pair(big)
    a.encode_steps_noalloc(60)
    b.encode_steps(60)
    n1 = _same(a, b)
    assert n1 == 60
    # second call, smaller input: a D2Q9 tunnel with a small cylinder (fewer partials than the first
    # call wrote), then the first handle again with a shorter run
    cfg2, sdf_fn = CASES2["tunnel"]

    def small():
        s = Solver.create(dict(cfg2))
        s.set_sdf(sdf_fn(cfg2["width"], cfg2["height"]))
        s.init_field()
        return s

    c, d = _pair(small)
    c.encode_steps_noalloc(25)
    d.encode_steps(25)
    assert _same(c, d) == 25
    a.encode_steps_noalloc(7)
    b.encode_steps(7)
    assert _same(a, b) == 7
    for s in (a, b, c, d):
        s.destroy()


def test_noalloc_refuses_reductions_beyond_the_static_scratch():
    # D3Q27 reduces over every wave of the grid: 256 x 128 x 128 needs 65,536 partials
    cfg = dict(width=256, height=128, depth=128, precision="fp16", tau=0.55, uIn=0.05, velocitySet=27)
    s = Solver3D.create(cfg)
    s.init_field()
    with pytest.raises(LbmError) as e:
        s.encode_steps_noalloc(1)
    assert "65536 partials" in str(e.value) and "32768" in str(e.value)
    assert s.step == 0
    s.destroy()
