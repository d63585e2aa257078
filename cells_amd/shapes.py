"""Parametrized obstacle SDFs for shape gradients (no counterpart in the reference).

Both return the float32 SDF the solver takes through set_sdf and, in float64, its derivative with
respect to every parameter at every node, which the adjoint combines with dJ/dq of every Bouzidi
link (q depends on the SDF at the link's two end nodes, src/shaders/step.ts:140-142).

* NACA0012 with thickness bumps: the SDF of the section (src/geometry.ts:42-80, polygon of 400
  points per surface) pitched nose-up by alpha degrees about its leading edge, minus
  sum_k a_k exp(-(xi - xi_k)^2 / (2 sigma^2)), where xi is the node's coordinate along the chord
  direction (cos alpha, -sin alpha) from the leading edge, xi_k = (k + 1/2) chord / K (centres
  uniform along the chord) and sigma = chord / 16. A positive a_k thickens the section near xi_k.
  Parameters: [alpha (degrees), a_1 .. a_K]. d/dalpha is a central difference of the polygon SDF
  in double precision with a 1e-6 degree step; d/da_k is analytic.
* Sphere with a spherical-harmonic radius: |x - c| - (r + sum_lm b_lm Y_lm(x_hat)) with the real,
  orthonormal spherical harmonics up to degree 3 (16 functions, ordered l = 0..3, m = -l..l).
  d/db_lm = -Y_lm(x_hat), analytic.
Values are computed inside the same stamp box as the reference's shapes (bounding box + 3 cells);
elsewhere the SDF stays FAR and the derivatives are zero.
"""
from __future__ import annotations

import math

import numpy as np

from . import geometry as g


def _box(minx, miny, maxx, maxy, W, H):
    x0 = max(0, math.floor(minx - g.MARGIN))
    y0 = max(0, math.floor(miny - g.MARGIN))
    x1 = min(W - 1, math.ceil(maxx + g.MARGIN))
    y1 = min(H - 1, math.ceil(maxy + g.MARGIN))
    return x0, y0, x1, y1


def naca_bumps(W, H, x0, y0, chord, alpha_deg, amps, sigma=None, n=400):
    """Returns (sdf float32 [H*W], dsdf float64 [1 + K, H*W]) for parameters [alpha_deg, amps...]."""
    amps = np.asarray(amps, dtype=np.float64)
    K = len(amps)
    sigma = chord / 16 if sigma is None else sigma
    a = math.radians(alpha_deg)
    # the stamp box of the unperturbed section at this angle, grown a little so that small bumps
    # and angle changes keep it
    poly = g.naca_polygon(x0, y0, chord, 0.12, a, n)
    P = np.asarray(poly)
    bx0, by0, bx1, by1 = _box(P[:, 0].min() - 1, P[:, 1].min() - 1, P[:, 0].max() + 1, P[:, 1].max() + 1, W, H)
    ys, xs = np.mgrid[by0 : by1 + 1, bx0 : bx1 + 1].astype(np.float64)

    def base(alpha):
        return g.polygon_distance(g.naca_polygon(x0, y0, chord, 0.12, math.radians(alpha), n), xs, ys)

    def xi(alpha):
        ar = math.radians(alpha)
        return (xs - x0) * math.cos(ar) - (ys - y0) * math.sin(ar)

    centres = (np.arange(K) + 0.5) * chord / K
    xv = xi(alpha_deg)
    G = np.exp(-((xv[None] - centres[:, None, None]) ** 2) / (2 * sigma * sigma))  # [K, h, w]
    val = base(alpha_deg) - np.tensordot(amps, G, axes=(0, 0))
    d = 1e-6

    def full(alpha):
        return base(alpha) - np.tensordot(amps, np.exp(-((xi(alpha)[None] - centres[:, None, None]) ** 2)
                                                       / (2 * sigma * sigma)), axes=(0, 0))

    dalpha = (full(alpha_deg + d) - full(alpha_deg - d)) / (2 * d)
    sdf = g.empty_sdf(W, H)
    view = sdf.reshape(H, W)[by0 : by1 + 1, bx0 : bx1 + 1]
    view[...] = np.minimum(view.astype(np.float64), val).astype(np.float32)
    grad = np.zeros((1 + K, H, W))
    grad[0, by0 : by1 + 1, bx0 : bx1 + 1] = dalpha
    grad[1:, by0 : by1 + 1, bx0 : bx1 + 1] = -G
    return sdf, grad.reshape(1 + K, H * W)


def real_sh(lmax, x, y, z):
    """Real orthonormal spherical harmonics Y_lm (l <= lmax, m = -l..l) at unit vectors (x, y, z),
    from the associated Legendre recurrences; returns [(lmax + 1)^2, ...]."""
    ct = z
    st = np.sqrt(np.maximum(0.0, 1 - z * z))
    phi = np.arctan2(y, x)
    out = []
    # associated Legendre P_l^m(ct) without the Condon-Shortley phase
    P = {}
    P[(0, 0)] = np.ones_like(ct)
    for m in range(1, lmax + 1):
        P[(m, m)] = (2 * m - 1) * st * P[(m - 1, m - 1)]
    for m in range(0, lmax):
        P[(m + 1, m)] = (2 * m + 1) * ct * P[(m, m)]
    for m in range(0, lmax + 1):
        for l in range(m + 2, lmax + 1):
            P[(l, m)] = ((2 * l - 1) * ct * P[(l - 1, m)] - (l + m - 1) * P[(l - 2, m)]) / (l - m)
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            am = abs(m)
            K = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - am) / math.factorial(l + am))
            if m == 0:
                out.append(K * P[(l, 0)])
            elif m > 0:
                out.append(math.sqrt(2) * K * P[(l, am)] * np.cos(am * phi))
            else:
                out.append(math.sqrt(2) * K * P[(l, am)] * np.sin(am * phi))
    return np.stack(out)


def sphere_sh(W, H, D, c, r, coeffs, lmax=3):
    """Returns (sdf float32 [N], dsdf float64 [(lmax+1)^2, N]) for the radius r + sum b_lm Y_lm."""
    coeffs = np.asarray(coeffs, dtype=np.float64)
    nb = (lmax + 1) ** 2
    assert coeffs.shape == (nb,)
    reach = r + np.abs(coeffs).sum() * 0.5 + 1
    lo = [max(0, math.floor(c[k] - reach - g.MARGIN)) for k in range(3)]
    hi = [min(n - 1, math.ceil(c[k] + reach + g.MARGIN)) for k, n in enumerate((W, H, D))]
    zs, ys, xs = np.mgrid[lo[2] : hi[2] + 1, lo[1] : hi[1] + 1, lo[0] : hi[0] + 1].astype(np.float64)
    dx, dy, dz = xs - c[0], ys - c[1], zs - c[2]
    rr = np.sqrt(dx * dx + dy * dy + dz * dz)
    safe = np.maximum(rr, 1e-12)
    Y = real_sh(lmax, dx / safe, dy / safe, dz / safe)
    val = rr - (r + np.tensordot(coeffs, Y, axes=(0, 0)))
    sdf = g.empty_sdf3(W, H, D)
    view = sdf.reshape(D, H, W)[lo[2] : hi[2] + 1, lo[1] : hi[1] + 1, lo[0] : hi[0] + 1]
    view[...] = np.minimum(view.astype(np.float64), val).astype(np.float32)
    grad = np.zeros((nb, D, H, W))
    grad[:, lo[2] : hi[2] + 1, lo[1] : hi[1] + 1, lo[0] : hi[0] + 1] = -Y
    return sdf, grad.reshape(nb, -1)
