"""Shared inputs for the GPU-vs-oracle parity tests: small tunnels that exercise every boundary rule."""
from __future__ import annotations

import numpy as np

from cells_amd import geometry as g
from cells_amd.lbm import X_OPEN, X_PERIODIC, X_WALL, Y_MOVING, Y_NOSLIP, Y_PERIODIC, Y_SLIP

# 2D: (name, config, sdf builder)
CASES2 = {
    # Config-0 shape in small: Zou-He inlet/outlet, slip walls, sponge + absorbing layer, circle.
    "tunnel": (dict(width=64, height=24, left=X_OPEN, right=X_OPEN, bottom=Y_SLIP, top=Y_SLIP, tau=0.548,
                    uIn=0.06, spongeFraction=0.15, spongeTau=1, absorb=0.02),
               lambda W, H: _circle(W, H, 16, 11.5, 4)),
    # Schaefer-Turek style: parabolic inlet, no-slip walls, inlet absorbing layer, Smagorinsky.
    "channel": (dict(width=56, height=20, left=X_OPEN, right=X_OPEN, bottom=Y_NOSLIP, top=Y_NOSLIP, tau=0.53,
                     uIn=0.05, parabolic=True, spongeFraction=0.1, spongeTau=1, absorb=0.1,
                     inletLayerFraction=0.08, smagorinsky=0.1),
                lambda W, H: _circle(W, H, 12, 9.5, 3.3)),
    # Periodic box with body force and a body across the seam (periodic wrap + Bouzidi).
    "periodic": (dict(width=40, height=20, left=X_PERIODIC, right=X_PERIODIC, bottom=Y_NOSLIP, top=Y_NOSLIP,
                      tau=0.8, bodyForce=(2e-5, 1e-6)),
                 lambda W, H: _circle(W, H, 1.5, 9.5, 3.0)),
    # Lid-driven cavity: x walls + moving lid.
    "cavity": (dict(width=24, height=24, left=X_WALL, right=X_WALL, bottom=Y_NOSLIP, top=Y_MOVING, tau=0.6,
                    uLid=0.08),
               None),
    # Fully periodic Taylor-Green-like start field.
    "box": (dict(width=32, height=32, left=X_PERIODIC, right=X_PERIODIC, bottom=Y_PERIODIC, top=Y_PERIODIC,
                 tau=0.8),
            None),
}


def _circle(W, H, cx, cy, r):
    s = g.empty_sdf(W, H)
    g.add_circle(s, W, H, cx, cy, r)
    return s


def taylor_green_macro(n, u0=0.04, t=0.0, nu=0.1):
    """Macro field (rho, ux, uy, 0) of the decaying Taylor-Green vortex (src/cases.ts:141-152)."""
    k = 2 * np.pi / n
    a = u0 * np.exp(-2 * nu * k * k * t)
    y, x = np.mgrid[0:n, 0:n].astype(np.float64)
    m = np.zeros((n, n, 4))
    m[..., 0] = 1 - 0.75 * a * a * (np.cos(2 * k * x) + np.cos(2 * k * y))
    m[..., 1] = -a * np.cos(k * x) * np.sin(k * y)
    m[..., 2] = a * np.sin(k * x) * np.cos(k * y)
    return m.reshape(-1).astype(np.float32)


# 3D: small tunnel with a sphere next to a slip face so every pullSlow3 branch fires.
def case3(precision="fp32", W=28, H=14, D=12, smag=0.0):
    cfg = dict(width=W, height=H, depth=D, precision=precision, tau=0.524, uIn=0.06, spongeFraction=0.15,
               spongeTau=1, absorb=0.02, smagorinsky=smag)
    sdf = g.empty_sdf3(W, H, D)
    g.add_sphere(sdf, W, H, D, 8, 6.5, 5.5, 3.2)
    # a second body touching the z = 0 face (slip face through a solid node)
    g.add_sphere(sdf, W, H, D, 17, 4, 0.5, 2.2)
    return cfg, sdf
