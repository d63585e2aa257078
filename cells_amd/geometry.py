"""Signed distance fields on the node grid, negative inside solids.

Host-side input producers for ``set_sdf``: the same shapes, margins and float32 rounding as the
reference's CPU geometry (src/geometry.ts:1-80, src/geometry3d.ts:1-99). Node (x, y[, z]) sits at
integer coordinates; the 2D index is ``y * W + x`` and the 3D index is ``x + W * (y + H * z)``.
Distances are computed in float64 and stored as float32, as the reference's Float32Array does.
"""
from __future__ import annotations

import math

import numpy as np

FAR = 1e6  # src/geometry.ts:3
MARGIN = 3  # src/geometry.ts:21, src/geometry3d.ts:6


def empty_sdf(W: int, H: int) -> np.ndarray:
    """emptySdf (src/geometry.ts:5-7)."""
    return np.full(W * H, FAR, dtype=np.float32)


def _stamp(sdf, W, H, box, d):
    """stamp (src/geometry.ts:9-19): min-union of d over the clipped integer box."""
    x0 = max(0, math.floor(box[0]))
    y0 = max(0, math.floor(box[1]))
    x1 = min(W - 1, math.ceil(box[2]))
    y1 = min(H - 1, math.ceil(box[3]))
    if x1 < x0 or y1 < y0:
        return
    ys, xs = np.mgrid[y0 : y1 + 1, x0 : x1 + 1].astype(np.float64)
    vals = d(xs, ys)
    view = sdf.reshape(H, W)[y0 : y1 + 1, x0 : x1 + 1]
    view[...] = np.minimum(view.astype(np.float64), vals).astype(np.float32)


def add_circle(sdf, W, H, cx, cy, r):
    """addCircle (src/geometry.ts:23-25)."""
    _stamp(sdf, W, H, [cx - r - MARGIN, cy - r - MARGIN, cx + r + MARGIN, cy + r + MARGIN],
           lambda x, y: np.hypot(x - cx, y - cy) - r)


def add_box(sdf, W, H, cx, cy, hx, hy, angle=0.0):
    """addBox (src/geometry.ts:28-39): rectangle of half-extents (hx, hy) rotated counter-clockwise."""
    c, s = math.cos(angle), math.sin(angle)
    R = math.hypot(hx, hy) + MARGIN

    def d(x, y):
        px = (x - cx) * c + (y - cy) * s
        py = -(x - cx) * s + (y - cy) * c
        qx = np.abs(px) - hx
        qy = np.abs(py) - hy
        return np.hypot(np.maximum(qx, 0), np.maximum(qy, 0)) + np.minimum(np.maximum(qx, qy), 0)

    _stamp(sdf, W, H, [cx - R, cy - R, cx + R, cy + R], d)


def naca_polygon(x0, y0, chord, thickness, alpha, n=400):
    """nacaPolygon (src/geometry.ts:42-57): symmetric 4-digit section, closed trailing edge,
    cosine spacing, leading edge at (x0, y0), pitched nose-up by alpha (radians)."""

    def half(t):
        return 5 * thickness * (0.2969 * math.sqrt(t) - 0.126 * t - 0.3516 * t * t + 0.2843 * t ** 3 - 0.1036 * t ** 4)

    pts = []
    for k in range(n + 1):
        t = (1 - math.cos(math.pi * k / n)) / 2
        pts.append((t, half(t)))
    for k in range(n - 1, 0, -1):
        t = (1 - math.cos(math.pi * k / n)) / 2
        pts.append((t, -half(t)))
    c, s = math.cos(-alpha), math.sin(-alpha)
    return [(x0 + chord * (u * c - v * s), y0 + chord * (u * s + v * c)) for u, v in pts]


def polygon_distance(poly, x, y):
    """Signed distance to a closed polygon (src/geometry.ts:65-78, src/geometry3d.ts:53-67):
    nearest-edge distance, negative where the even-odd crossing test says inside."""
    P = np.asarray(poly, dtype=np.float64)
    A = P
    B = np.roll(P, 1, axis=0)  # edge from poly[a] to poly[b = a - 1]
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    d2 = np.full(x.shape, np.inf)
    inside = np.zeros(x.shape, dtype=bool)
    for (ax, ay), (bx, by) in zip(A, B):
        ex, ey = bx - ax, by - ay
        t = np.clip(((x - ax) * ex + (y - ay) * ey) / (ex * ex + ey * ey), 0, 1)
        dx = x - ax - t * ex
        dy = y - ay - t * ey
        d2 = np.minimum(d2, dx * dx + dy * dy)
        crosses = (ay > y) != (by > y)
        if by != ay:
            inside ^= crosses & (x < (bx - ax) * (y - ay) / (by - ay) + ax)
    d = np.sqrt(d2)
    return np.where(inside, -d, d)


def add_polygon(sdf, W, H, poly):
    """addPolygon (src/geometry.ts:59-80)."""
    P = np.asarray(poly, dtype=np.float64)
    minx, miny = P.min(axis=0)
    maxx, maxy = P.max(axis=0)
    _stamp(sdf, W, H, [minx - MARGIN, miny - MARGIN, maxx + MARGIN, maxy + MARGIN],
           lambda x, y: polygon_distance(poly, x, y))


# ---------------------------------------------------------------- 3D (src/geometry3d.ts)

def empty_sdf3(W: int, H: int, D: int) -> np.ndarray:
    """emptySdf3 (src/geometry3d.ts:8-10)."""
    return np.full(W * H * D, FAR, dtype=np.float32)


def _stamp3(sdf, W, H, D, box, d):
    """stamp3 (src/geometry3d.ts:14-27): the box is grown by MARGIN, then clipped."""
    x0 = max(0, math.floor(box[0] - MARGIN))
    y0 = max(0, math.floor(box[1] - MARGIN))
    z0 = max(0, math.floor(box[2] - MARGIN))
    x1 = min(W - 1, math.ceil(box[3] + MARGIN))
    y1 = min(H - 1, math.ceil(box[4] + MARGIN))
    z1 = min(D - 1, math.ceil(box[5] + MARGIN))
    if x1 < x0 or y1 < y0 or z1 < z0:
        return
    zs, ys, xs = np.mgrid[z0 : z1 + 1, y0 : y1 + 1, x0 : x1 + 1].astype(np.float64)
    vals = d(xs, ys, zs)
    view = sdf.reshape(D, H, W)[z0 : z1 + 1, y0 : y1 + 1, x0 : x1 + 1]
    view[...] = np.minimum(view.astype(np.float64), vals).astype(np.float32)


def add_sphere(sdf, W, H, D, cx, cy, cz, r):
    """addSphere (src/geometry3d.ts:29-31)."""
    _stamp3(sdf, W, H, D, [cx - r, cy - r, cz - r, cx + r, cy + r, cz + r],
            lambda x, y, z: np.sqrt((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2) - r)


def add_box3(sdf, W, H, D, cx, cy, cz, hx, hy, hz, angle_z=0.0):
    """addBox3 (src/geometry3d.ts:34-46)."""
    c, s = math.cos(angle_z), math.sin(angle_z)
    R = math.hypot(hx, hy)

    def d(x, y, z):
        px = (x - cx) * c + (y - cy) * s
        py = -(x - cx) * s + (y - cy) * c
        qx = np.abs(px) - hx
        qy = np.abs(py) - hy
        qz = np.abs(z - cz) - hz
        out = np.sqrt(np.maximum(qx, 0) ** 2 + np.maximum(qy, 0) ** 2 + np.maximum(qz, 0) ** 2)
        return out + np.minimum(np.maximum(np.maximum(qx, qy), qz), 0)

    _stamp3(sdf, W, H, D, [cx - R, cy - R, cz - hz, cx + R, cy + R, cz + hz], d)


def add_cylinder_z(sdf, W, H, D, cx, cy, r):
    """addCylinderZ (src/geometry3d.ts:49-51)."""
    _stamp3(sdf, W, H, D, [cx - r, cy - r, 0, cx + r, cy + r, D - 1],
            lambda x, y, z: np.hypot(x - cx, y - cy) - r)


def add_wing(sdf, W, H, D, cx, cy, cz, chord, span, angle_z=0.0):
    """addWing (src/geometry3d.ts:70-88): NACA0012 section extruded over span, flat tips,
    pitched nose-up about mid-chord."""
    ox = cx - (chord / 2) * math.cos(angle_z)
    oy = cy + (chord / 2) * math.sin(angle_z)
    poly = [(x + ox, y + oy) for x, y in naca_polygon(0, 0, chord, 0.12, angle_z, 200)]
    P = np.asarray(poly)
    minx, miny = P.min(axis=0)
    maxx, maxy = P.max(axis=0)

    def d(x, y, z):
        d2 = polygon_distance(poly, x, y)
        dz = np.abs(z - cz) - span / 2
        return np.where((d2 < 0) | (dz < 0), np.maximum(d2, dz), np.hypot(d2, dz))

    _stamp3(sdf, W, H, D, [minx, miny, cz - span / 2, maxx, maxy, cz + span / 2], d)


def reference_area(obstacle: str, size: float, D: int) -> float:
    """referenceArea (src/geometry3d.ts:91-99): frontal area in cells^2."""
    return {
        "sphere": math.pi * size * size / 4,
        "cube": size * size,
        "cylinder": size * D,
        "wing": size * 0.6 * D,
        "none": 0.0,
    }[obstacle]
