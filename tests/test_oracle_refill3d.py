"""Pins the oracle's 3D geometry rebuild with refill (src/solver3d.ts:211-228, src/shaders/aux3d.ts:42-100)
and its D3Q19 / D3Q27 macro field (src/shaders/aux3d.ts:137-150) against what the definitions fix.

A cell that was solid before setSdf and is fluid after restarts at the equilibrium of the mean density
and mean velocity of its in-grid neighbours x + c_i (i >= 1) that were fluid before the rebuild, or at
(rho, u) = (1, (uIn, 0, 0)) when it has none; every other cell keeps its populations. The expected
moments are computed here from the populations read before the rebuild (rho = 1 + sum(f - w),
u = sum c (f - w) / rho), not from the oracle's macro kernel, which is checked separately against the
same numpy moments, including rho = 0 and u = 0 in solids and the (ux, uy, uz, rho) layout."""
from __future__ import annotations

import numpy as np
import pytest

from cells_amd import geometry as g
from oracle import oracle as O

W, H, D = 30, 14, 13
UIN = 0.07


def _tables(q):
    t = O.tables27() if q == 27 else O.tables3()
    return t["cx"], t["cy"], t["cz"], t["w"]


def _solver(q, bits=64):
    cfg = dict(width=W, height=H, depth=D, precision="fp32", tau=0.56, uIn=UIN, spongeFraction=0.15, spongeTau=1,
               absorb=0.02, velocitySet=q)
    return O.OracleSolver3D(cfg, bits)


def _moments(f, q):
    """(rho, ux, uy, uz) per cell from f - w of shape (Q, N), in float64."""
    cx, cy, cz, _ = _tables(q)
    rho = 1.0 + f.sum(0)
    return rho, (cx @ f) / rho, (cy @ f) / rho, (cz @ f) / rho


def _sdf_before():
    sdf = g.empty_sdf3(W, H, D)
    g.add_sphere(sdf, W, H, D, 11.0, 6.5, 6.0, 4.6)
    return sdf


def _sdf_after():
    """The sphere moved 2 cells downstream (its upstream cap uncovered) with a small cavity carved at
    the old centre, which lies deep in the old body: that cavity's cells have no fluid neighbour."""
    sdf = g.empty_sdf3(W, H, D)
    g.add_sphere(sdf, W, H, D, 13.0, 6.5, 6.0, 4.6)
    z, y, x = np.mgrid[0:D, 0:H, 0:W].astype(np.float64)
    cavity = np.sqrt((x - 11.0) ** 2 + (y - 6.5) ** 2 + (z - 6.0) ** 2) - 0.9
    return np.maximum(sdf.reshape(D, H, W), -cavity).astype(np.float32).reshape(-1)


@pytest.mark.parametrize("q", [19, 27])
def test_refill_takes_mean_of_formerly_fluid_neighbours(q):
    cx, cy, cz, w = _tables(q)
    solid_bit = np.uint32(O.flag_bits(q)[0])
    s = _solver(q)
    s.set_sdf(_sdf_before())
    s.init_field()
    s.steps(50)
    f0 = s.read_populations()
    rho0, ux0, uy0, uz0 = _moments(f0, q)
    solid0 = (s.read_flags() & solid_bit) != 0
    s.set_sdf(_sdf_after())
    solid1 = (s.read_flags() & solid_bit) != 0
    f1 = s.read_populations()
    fluid0 = ~solid0.reshape(D, H, W)
    uncovered = np.flatnonzero(solid0 & ~solid1)
    assert len(uncovered) > 20
    # cells that were fluid before keep their populations bit for bit (solid ones are not read)
    kept = ~solid0 & ~solid1
    np.testing.assert_array_equal(f1[:, kept], f0[:, kept])
    seen_rest = seen_mean = 0
    ful = f1 + w[:, None]
    for k in uncovered:
        z, y, x = int(k) // (W * H), (int(k) // W) % H, int(k) % W
        nb = [(z + c, y + b, x + a) for a, b, c in zip(cx[1:], cy[1:], cz[1:])
              if 0 <= x + a < W and 0 <= y + b < H and 0 <= z + c < D and fluid0[z + c, y + b, x + a]]
        if nb:
            ids = [(zz * H + yy) * W + xx for zz, yy, xx in nb]
            rho, u = rho0[ids].mean(), (ux0[ids].mean(), uy0[ids].mean(), uz0[ids].mean())
            seen_mean += 1
        else:
            rho, u = 1.0, (UIN, 0.0, 0.0)
            seen_rest += 1
        fc = ful[:, k]
        assert abs(fc.sum() - rho) < 1e-12, (k, fc.sum(), rho)
        for c, ua in zip((cx, cy, cz), u):
            assert abs((fc * c).sum() - rho * ua) < 1e-12
        # equilibrium second moments: rho u_a u_b off the diagonal, rho / 3 + rho u_a^2 on it
        assert abs((fc * cx * cy).sum() - rho * u[0] * u[1]) < 1e-12
        assert abs((fc * cy * cz).sum() - rho * u[1] * u[2]) < 1e-12
        assert abs((fc * cx * cx).sum() - (rho / 3 + rho * u[0] ** 2)) < 1e-12
    assert seen_rest >= 1 and seen_mean >= 20


@pytest.mark.parametrize("q", [19, 27])
def test_macro_layout_and_solids(q):
    s = _solver(q)
    s.set_sdf(_sdf_before())
    s.init_field()
    s.steps(30)
    f = s.read_populations()
    rho, ux, uy, uz = _moments(f, q)
    m = s.macro().reshape(-1, 4)
    solid = (s.read_flags() & np.uint32(O.flag_bits(q)[0])) != 0
    assert solid.sum() > 100
    np.testing.assert_array_equal(m[solid], 0.0)
    fl = ~solid
    for col, ref in ((0, ux), (1, uy), (2, uz), (3, rho)):
        assert np.abs(m[fl, col] - ref[fl]).max() < 1e-14
    assert np.abs(m[fl, 0]).max() > 0.05 and np.abs(m[fl, 3] - 1).max() > 1e-4  # a developing flow
