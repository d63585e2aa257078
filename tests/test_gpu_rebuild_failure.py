"""A geometry rebuild that fails part way (forced through the lbm_debug_fail_rebuilds test hook, at
the point where the new boundary tables would be allocated): the handle must not step with device
flags that no longer match its tables, lbm_set_open_faces must fall back to the previous kind by
itself, and once a rebuild succeeds the tunnel must compute exactly what a tunnel that never saw the
failure computes."""
from __future__ import annotations

import numpy as np
import pytest

from cells_amd import geometry as g
from cells_amd.lbm import LbmError, Solver, Solver3D
from tests.parity_cases import CASES2, case3

pytestmark = pytest.mark.gpu


def _same(a, b):
    np.testing.assert_array_equal(a.read_populations_raw(), b.read_populations_raw())
    np.testing.assert_array_equal(a.read_flags(), b.read_flags())
    sa, va = a.read_forces_raw()
    sb, vb = b.read_forces_raw()
    np.testing.assert_array_equal(sa, sb)
    np.testing.assert_array_equal(va, vb)


@pytest.mark.parametrize("variant", [0, 4])
def test_open_faces_failure_keeps_previous_faces(variant):
    cfg, sdf = case3("fp32")
    W, H, D = cfg["width"], cfg["height"], cfg["depth"]
    g.add_sphere(sdf, W, H, D, 2.6, 8.0, 6.5, 1.9)  # inlet nodes with body links: faces change the list
    a, b = Solver3D.create(dict(cfg)), Solver3D.create(dict(cfg))
    for s in (a, b):
        s.set_sdf(sdf)
        s.init_field()
        s.set_launch(64, variant)
        s.run(15)
    a.debug_fail_rebuilds(1)
    with pytest.raises(LbmError) as e:
        a.set_open_faces("zou-he")
    assert "test hook" in str(e.value) and "previous faces are kept" in str(e.value)
    for s in (a, b):
        s.run(20)
    _same(a, b)
    # both the switch and the fallback fail: stepping is refused until a rebuild succeeds
    a.debug_fail_rebuilds(2)
    with pytest.raises(LbmError) as e:
        a.set_open_faces("zou-he")
    assert "failed too" in str(e.value)
    step = a.step
    with pytest.raises(LbmError):
        a.run(1)
    with pytest.raises(LbmError):
        a.time_steps(1)
    assert a.step == step
    a.rebuild_flags()
    for s in (a, b):
        s.run(20)
    _same(a, b)
    # the switch itself still works afterwards
    for s in (a, b):
        s.set_open_faces("zou-he")
        s.run(10)
    _same(a, b)
    a.destroy()
    b.destroy()


def test_set_sdf_failure_refuses_steps_until_rebuilt():
    name = "tunnel"
    cfg, sdf_fn = CASES2[name]
    W, H = cfg["width"], cfg["height"]
    a, b = Solver.create(dict(cfg)), Solver.create(dict(cfg))
    for s in (a, b):
        s.set_sdf(sdf_fn(W, H))
        s.init_field()
        s.run(30)
    sdf2 = g.empty_sdf(W, H)
    g.add_circle(sdf2, W, H, 20, 11.5, 3)  # moves the body: cells are uncovered and refilled
    a.debug_fail_rebuilds(1)
    with pytest.raises(LbmError) as e:
        a.set_sdf(sdf2)
    assert "test hook" in str(e.value)
    with pytest.raises(LbmError) as e:
        a.run(1)
    assert "geometry rebuild failed" in str(e.value)
    a.set_sdf(sdf2)
    b.set_sdf(sdf2)
    for s in (a, b):
        s.run(25)
    _same(a, b)
    a.destroy()
    b.destroy()
