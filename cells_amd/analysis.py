"""Units, force coefficients and signal analysis on the host.

Restates src/units.ts:1-24, src/analysis.ts:1-55 and the small helpers of src/tunnel3d.ts:75-105
that turn force samples into C_D, C_L and the Strouhal number.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

TAU_MIN = 0.51  # src/units.ts:1


@dataclass
class Derived:
    nu: float
    tau: float
    clamped: bool
    effective_re: float
    mach: float


def derive_tau(re: float, u_lat: float, length: float) -> Derived:
    """deriveTau (src/units.ts:14-20): tau = 3 nu + 1/2 with nu = u L / Re, clamped at 0.51."""
    nu = u_lat * length / re
    raw = 3 * nu + 0.5
    tau = max(raw, TAU_MIN)
    nu_eff = (tau - 0.5) / 3
    return Derived(nu_eff, tau, raw < TAU_MIN, u_lat * length / nu_eff, u_lat * math.sqrt(3))


def max_reynolds(u_lat: float, length: float) -> float:
    """maxReynolds (src/units.ts:22-24)."""
    return u_lat * length * 3 / (TAU_MIN - 0.5)


def period_from_crossings(t, v):
    """periodFromCrossings (src/analysis.ts:8-18): mean spacing of linearly interpolated upward
    crossings of the signal's mean; None with fewer than 3 crossings.
    Returns (period, crossings)."""
    v = [float(a) for a in v]
    t = [float(a) for a in t]
    m = sum(v) / len(v)
    times = []
    for k in range(1, len(v)):
        a = v[k - 1] - m
        b = v[k] - m
        if a < 0 and b >= 0:
            times.append(t[k - 1] + (t[k] - t[k - 1]) * -a / (b - a))
    if len(times) < 3:
        return None
    return (times[-1] - times[0]) / (len(times) - 1), len(times)


def mean(v) -> float:
    """mean (src/analysis.ts:20-22)."""
    v = list(v)
    return sum(v) / len(v)


def stream_function(ux, uy, W: int, H: int) -> np.ndarray:
    """streamFunction (src/analysis.ts:28-37), psi[y * W + x]."""
    ux = np.asarray(ux, dtype=np.float64)
    uy = np.asarray(uy, dtype=np.float64)
    psi = np.zeros(W * H)
    for y in range(1, H):
        psi[y * W] = psi[(y - 1) * W] + 0.5 * (ux[y * W] + ux[(y - 1) * W])
    for y in range(H):
        for x in range(1, W):
            i = y * W + x
            psi[i] = psi[i - 1] - 0.5 * (uy[i] + uy[i - 1])
    return psi


def primary_vortex(psi, W: int, H: int):
    """primaryVortex (src/analysis.ts:40-55): extremum of |psi| refined by a quadratic fit."""
    best = 1
    for y in range(1, H - 1):
        for x in range(1, W - 1):
            if abs(psi[y * W + x]) > abs(psi[best]):
                best = y * W + x
    bx, by = best % W, best // W

    def refine(m, c, p):
        den = m - 2 * c + p
        return 0.0 if den == 0 else 0.5 * (m - p) / den

    return (bx + refine(psi[best - 1], psi[best], psi[best + 1]),
            by + refine(psi[best - W], psi[best], psi[best + W]), psi[best])


def coefficient_scale(u_target: float, area: float) -> float:
    """coefficientScale (src/tunnel3d.ts:75-77): 2 / (u^2 A), 0 without a body."""
    return 2 / (u_target * u_target * area) if area > 0 else 0.0


def has_diverged(samples) -> bool:
    """hasDiverged (src/tunnel3d.ts:97-99): any non-finite force component."""
    return any(not math.isfinite(c) for s in samples for c in s[1:])


def field_diverged(values) -> bool:
    """fieldDiverged (src/tunnel3d.ts:102-105): any non-finite macro value."""
    return not bool(np.isfinite(np.asarray(values)).all())


def smoothstep_ramp(t: float) -> float:
    """Inlet ramp of the reference drivers (src/cases.ts:52-53, src/cases3d.ts:9): t^2 (3 - 2t)."""
    return t * t * (3 - 2 * t)
