"""Pins the oracle's brush (src/solver.ts:243-268, src/shaders/aux.ts:176-196) against what the
geometry fixes: the written box is the disc's box grown by 2 cells and clipped to the grid, the
values in it are the distance to the disc in float (checked against float64), union is a min and
erase a max with the negated distance, a stroke is its discs in order followed by one rebuild, and
the rebuilt flags equal those of setSdf on the same field."""
from __future__ import annotations

import math

import numpy as np
import pytest

from cells_amd.lbm import X_OPEN, Y_SLIP
from oracle import oracle as O

FAR = 1e6


def tunnel(W=40, H=20):
    return O.OracleSolver(dict(width=W, height=H, left=X_OPEN, right=X_OPEN, bottom=Y_SLIP, top=Y_SLIP, tau=0.6,
                               uIn=0.04, forceEvery=1), 32)


def changed_box(sdf, W, H):
    ys, xs = np.nonzero(sdf.reshape(H, W) != np.float32(FAR))
    return (xs.min(), ys.min(), xs.max(), ys.max()) if xs.size else None


def exact_distance(W, H, cx, cy, r):
    y, x = np.mgrid[0:H, 0:W].astype(np.float64)
    return np.hypot(x - np.float32(cx), y - np.float32(cy)) - np.float32(r)


@pytest.mark.parametrize("cx,cy,r", [(10.3, 7.6, 3.0), (20.0, 10.0, 2.5), (5.5, 9.25, 4.0)])
def test_union_box_and_values(cx, cy, r):
    W, H = 40, 20
    s = tunnel(W, H)
    s.brush(cx, cy, r)
    sdf = s.read_sdf()
    box = changed_box(sdf, W, H)
    # every cell of [floor(c - r - 2), ceil(c + r + 2)] inside the grid is written, none outside it
    assert box == (max(0, math.floor(cx - r - 2)), max(0, math.floor(cy - r - 2)),
                   min(W - 1, math.ceil(cx + r + 2)), min(H - 1, math.ceil(cy + r + 2)))
    x0, y0, x1, y1 = box
    d = exact_distance(W, H, cx, cy, r)[y0:y1 + 1, x0:x1 + 1]
    got = sdf.reshape(H, W)[y0:y1 + 1, x0:x1 + 1].astype(np.float64)
    assert np.max(np.abs(got - d)) < 4e-6 * (1 + np.max(np.abs(d)))
    # cells strictly inside the disc are solid in the rebuilt flags, cells outside it are not
    solid = (s.read_flags() & np.uint32(O.flag_bits(2)[0])) != 0
    inside = (exact_distance(W, H, cx, cy, r) < -1e-5).reshape(-1)
    outside = (exact_distance(W, H, cx, cy, r) > 1e-5).reshape(-1)
    assert inside.any() and np.all(solid[inside]) and not np.any(solid[outside])


def test_clipping_and_empty_boxes():
    W, H = 40, 20
    s = tunnel(W, H)
    s.brush(0.5, 0.5, 3.0)
    assert changed_box(s.read_sdf(), W, H) == (0, 0, 6, 6)
    s.brush(38.2, 18.9, 1.0)
    assert changed_box(s.read_sdf(), W, H) == (0, 0, 39, 19)
    t = tunnel(W, H)
    for cx, cy in [(-20.0, 10.0), (100.0, 10.0), (20.0, -30.0), (20.0, 60.0)]:
        t.brush(cx, cy, 3.0)
    assert np.all(t.read_sdf() == np.float32(FAR))
    # a box that reaches into the grid writes there, even for a disc centred outside it
    t.brush(-3.5, 10.0, 2.0)
    assert changed_box(t.read_sdf(), W, H) == (0, 6, 1, 14)


def test_erase_is_max_with_negated_distance():
    W, H = 40, 20
    s = tunnel(W, H)
    s.brush(20.0, 10.0, 6.0)
    before = s.read_sdf().astype(np.float64)
    s.brush(20.0, 10.0, 2.5, erase=True)
    after = s.read_sdf().astype(np.float64)
    d = exact_distance(W, H, 20.0, 10.0, 2.5).reshape(-1)
    box = np.zeros((H, W), bool)
    box[6:15, 16:25] = True  # [floor(20 - 4.5), ceil(20 + 4.5)] x [floor(10 - 4.5), ceil(10 + 4.5)]
    box = box.reshape(-1)
    np.testing.assert_array_equal(after[~box], before[~box])
    want = np.maximum(before[box], -d[box])
    assert np.max(np.abs(after[box] - want)) < 4e-6
    # a ring remains: solid between r = 2.5 and r = 6, fluid inside r = 2.5
    y, x = np.divmod(np.arange(W * H), W)
    rr = np.hypot(x - 20.0, y - 10.0)
    assert np.all(after[rr < 2.4] > 0) and np.all(after[(rr > 2.6) & (rr < 5.9)] < 0)


def test_stroke_equals_discs_then_one_rebuild_and_set_sdf():
    W, H = 40, 20
    pts = [(8.0, 5.0), (9.7, 6.1), (11.4, 7.2), (13.1, 8.3)]
    a, b, c = tunnel(W, H), tunnel(W, H), tunnel(W, H)
    for s in (a, b, c):
        s.init_field()
        s.steps(30)
    a.brush_stroke(pts, 2.0)
    for x, y in pts:
        b.brush(x, y, 2.0, rebuild=False)
    b.rebuild_flags()
    np.testing.assert_array_equal(a.read_sdf(), b.read_sdf())
    np.testing.assert_array_equal(a.read_flags(), b.read_flags())
    np.testing.assert_array_equal(a.read_populations(), b.read_populations())
    # the same field through setSdf: same flags and the same refill
    c.set_sdf(a.read_sdf())
    np.testing.assert_array_equal(a.read_flags(), c.read_flags())
    np.testing.assert_array_equal(a.read_populations(), c.read_populations())
    # the stroke's discs are a min-union: order does not matter for the field
    d = tunnel(W, H)
    d.brush_stroke(pts[::-1], 2.0)
    np.testing.assert_array_equal(a.read_sdf(), d.read_sdf())


def test_brush_without_rebuild_leaves_flags():
    W, H = 40, 20
    s = tunnel(W, H)
    f0 = s.read_flags()
    s.brush(20.0, 10.0, 4.0, rebuild=False)
    np.testing.assert_array_equal(s.read_flags(), f0)
    s.rebuild_flags()
    assert np.any(s.read_flags() != f0)


def test_erase_refills_uncovered_cells_from_fluid_neighbours():
    """src/shaders/aux.ts:74-87: a cell the body uncovered restarts at equilibrium with the mean
    density and mean velocity of its in-grid neighbours that were fluid before the rebuild, or at
    rest when it has none. Checked through the moments of the refilled populations (FP64 oracle)."""
    W, H = 40, 20
    s = O.OracleSolver(dict(width=W, height=H, left=X_OPEN, right=X_OPEN, bottom=Y_SLIP, top=Y_SLIP, tau=0.6,
                            uIn=0.05, forceEvery=1), 64)
    s.brush(14.0, 9.5, 5.0)
    s.init_field()
    s.steps(60)
    mac = s.macro().reshape(H, W, 4)
    solid0 = mac[..., 3] != 0
    s.brush(10.0, 9.5, 2.2, erase=True)  # bites into the body's upstream side
    s.brush(14.0, 9.5, 1.0, erase=True)  # a pocket with no fluid neighbour
    t = O.tables2()
    f = s.read_populations().reshape(9, H, W) + t["w"][:, None, None]
    solid1 = (s.read_flags().reshape(H, W) & np.uint32(O.flag_bits(2)[0])) != 0
    uncovered = np.argwhere(solid0 & ~solid1)
    assert len(uncovered) > 4
    seen_rest = seen_avg = 0
    for y, x in uncovered:
        nb = [(y + dy, x + dx) for dx, dy in zip(t["cx"][1:], t["cy"][1:])
              if 0 <= x + dx < W and 0 <= y + dy < H and not solid0[y + dy, x + dx]]
        rho = np.mean([mac[p][0] for p in nb]) if nb else 1.0
        ux = np.mean([mac[p][1] for p in nb]) if nb else 0.0
        uy = np.mean([mac[p][2] for p in nb]) if nb else 0.0
        fc = f[:, y, x]
        assert abs(fc.sum() - rho) < 1e-12
        assert abs((fc * t["cx"]).sum() - rho * ux) < 1e-12
        assert abs((fc * t["cy"]).sum() - rho * uy) < 1e-12
        assert abs((fc * t["cx"] * t["cy"]).sum() - rho * ux * uy) < 1e-12  # equilibrium stress
        seen_rest += not nb
        seen_avg += bool(nb)
    assert seen_rest >= 1 and seen_avg >= 4
