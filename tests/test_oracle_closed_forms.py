"""Pins the oracle's stream/collide/boundary arithmetic against closed forms and invariants.

Each test is chosen so that a plausible restatement slip (a dropped source term, a sign, a wrong
pair, a transposed index, a wrong Bouzidi branch) fails it:
  * plane Poiseuille with Guo forcing and halfway bounce-back, L2 < 2e-4 (src/cases.ts:61-92);
  * Taylor-Green decay: measured viscosity within 1%, L2 < 1e-2, 2nd-order convergence
    (src/cases.ts:127-189);
  * Bouzidi with q = 1/2 everywhere reduces to halfway bounce-back, and the momentum-exchange force
    on the walls balances the body force (steady state);
  * Zou-He inlet velocity and outlet density, the 3D equilibrium inlet, rest as a fixed point, mass
    conservation in a periodic box, the y/z mirror symmetry of the 3D sphere force.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from cells_amd import geometry as g
from cells_amd.lbm import X_OPEN, X_PERIODIC, Y_NOSLIP, Y_PERIODIC, Y_SLIP
from oracle import oracle as O


def _poiseuille_profile(tau, H=32, walls="noslip", steps=None):
    nu = (tau - 0.5) / 3
    u_max = 0.05
    fx = 8 * nu * u_max / (H * H)
    n = steps or math.ceil(10 * H * H / nu)
    if walls == "noslip":
        s = O.OracleSolver(dict(width=4, height=H, left=X_PERIODIC, right=X_PERIODIC, bottom=Y_NOSLIP,
                                top=Y_NOSLIP, tau=tau, bodyForce=(fx, 0), forceEvery=0), 64)
        s.init_field()
        s.steps(n)
        m = s.macro().reshape(H, 4, 4)
        return m[:, 1, 1], fx, nu, s
    # Bouzidi walls at exactly half a link from the first and last fluid rows: solid rows y = 0 and
    # y = H + 1 with the SDF planes at y = 0.5 and y = H + 0.5, so q = 1/2 on every wall link.
    Ht = H + 2
    s = O.OracleSolver(dict(width=4, height=Ht, left=X_PERIODIC, right=X_PERIODIC, bottom=Y_PERIODIC,
                            top=Y_PERIODIC, tau=tau, bodyForce=(fx, 0), forceEvery=1), 64)
    y = np.repeat(np.arange(Ht, dtype=np.float64), 4)
    sdf = np.minimum(y - 0.5, (H + 0.5) - y).astype(np.float32)
    s.set_sdf(sdf)
    s.init_field()
    s.steps(n)
    m = s.macro().reshape(Ht, 4, 4)
    return m[1 : H + 1, 1, 1], fx, nu, s


@pytest.mark.parametrize("tau", [0.6, 1.0, 1.5])
def test_poiseuille_exact_parabola(tau):
    H = 32
    ux, fx, nu, _ = _poiseuille_profile(tau, H)
    yy = np.arange(H) + 0.5
    exact = fx / (2 * nu) * yy * (H - yy)
    err = math.sqrt(((ux - exact) ** 2).sum() / (exact ** 2).sum())
    assert err < 2e-4, err


def test_poiseuille_wrong_lambda_fails():
    """The same test detects a wrong magic parameter (Lambda = 1/4 gives ~4.5e-4, src/cases.ts:86)."""
    H, tau = 32, 0.6
    nu = (tau - 0.5) / 3
    fx = 8 * nu * 0.05 / (H * H)
    s = O.OracleSolver(dict(width=4, height=H, left=X_PERIODIC, right=X_PERIODIC, bottom=Y_NOSLIP, top=Y_NOSLIP,
                            tau=tau, bodyForce=(fx, 0), forceEvery=0, **{"lambda": 0.25}), 64)
    s.init_field()
    s.steps(math.ceil(10 * H * H / nu))
    ux = s.macro().reshape(H, 4, 4)[:, 1, 1]
    yy = np.arange(H) + 0.5
    exact = fx / (2 * nu) * yy * (H - yy)
    err = math.sqrt(((ux - exact) ** 2).sum() / (exact ** 2).sum())
    assert err > 3e-4, err


def test_bouzidi_half_link_equals_halfway_bounce_back_and_force_balance():
    tau, H = 0.8, 16
    steps = 6000
    ux_ns, fx, nu, _ = _poiseuille_profile(tau, H, "noslip", steps)
    ux_bz, fx2, _, s = _poiseuille_profile(tau, H, "bouzidi", steps)
    np.testing.assert_allclose(ux_bz, ux_ns, rtol=0, atol=1e-13)
    # Steady momentum balance: the walls carry the body force of every fluid node.
    f = s.read_forces()
    fx_wall = f[-1].fx
    assert abs(fx_wall - fx * 4 * H) / (fx * 4 * H) < 1e-6, (fx_wall, fx * 4 * H)
    assert abs(f[-1].fy) < 1e-12


def _tg_run(n, tau=0.8):
    nu = (tau - 0.5) / 3
    u0 = 0.04 * 32 / n
    k = 2 * math.pi / n

    def field(t):
        a = u0 * math.exp(-2 * nu * k * k * t)
        y, x = np.mgrid[0:n, 0:n].astype(np.float64)
        m = np.zeros((n, n, 4))
        m[..., 0] = 1 - 0.75 * a * a * (np.cos(2 * k * x) + np.cos(2 * k * y))
        m[..., 1] = -a * np.cos(k * x) * np.sin(k * y)
        m[..., 2] = a * np.sin(k * x) * np.cos(k * y)
        return m.reshape(-1).astype(np.float32)

    s = O.OracleSolver(dict(width=n, height=n, left=X_PERIODIC, right=X_PERIODIC, bottom=Y_PERIODIC,
                            top=Y_PERIODIC, tau=tau, forceEvery=0), 64)
    s.init_field(field(0))
    t = round(math.log(2) / (2 * nu * k * k))
    s.steps(round(t / 2))
    mid = s.macro().reshape(-1, 4)
    s.steps(t - round(t / 2))
    end = s.macro().reshape(-1, 4)
    ex = field(t).reshape(-1, 4).astype(np.float64)
    err = math.sqrt(((end[:, 1:3] - ex[:, 1:3]) ** 2).sum() / (ex[:, 1:3] ** 2).sum())
    energy = lambda m: (m[:, 1] ** 2 + m[:, 2] ** 2).sum()
    dt = t - round(t / 2)
    nu_m = math.log(energy(mid) / energy(end)) / (4 * k * k * dt)
    return err, nu_m / nu


def test_taylor_green_decay_and_order():
    errs = []
    for n in (32, 64, 128):
        e, ratio = _tg_run(n)
        errs.append(e)
    assert 0.99 <= ratio <= 1.01, ratio
    assert errs[2] < 0.01
    order = math.log2(errs[0] / errs[2]) / 2
    assert order >= 1.8, (errs, order)


def test_periodic_box_conserves_mass_and_momentum():
    n = 24
    s = O.OracleSolver(dict(width=n, height=n, left=X_PERIODIC, right=X_PERIODIC, bottom=Y_PERIODIC, top=Y_PERIODIC,
                            tau=0.7, forceEvery=0), 64)
    rng = np.random.default_rng(3)
    m = np.zeros((n * n, 4))
    m[:, 0] = 1 + 0.01 * rng.standard_normal(n * n)
    m[:, 1:3] = 0.02 * rng.standard_normal((n * n, 2))
    s.init_field(m.reshape(-1).astype(np.float32))
    t = O.tables2()
    f0 = s.read_populations()
    s.steps(200)
    f1 = s.read_populations()
    assert abs(f1.sum() - f0.sum()) < 1e-11
    assert abs((f1 * t["cx"][:, None]).sum() - (f0 * t["cx"][:, None]).sum()) < 1e-11
    assert abs((f1 * t["cy"][:, None]).sum() - (f0 * t["cy"][:, None]).sum()) < 1e-11


def test_rest_is_a_fixed_point():
    s = O.OracleSolver(dict(width=20, height=10, left=X_OPEN, right=X_OPEN, bottom=Y_SLIP, top=Y_SLIP, tau=0.6,
                            spongeFraction=0.2, absorb=0.05, smagorinsky=0.1), 64)
    sdf = g.empty_sdf(20, 10)
    g.add_circle(sdf, 20, 10, 8, 4.5, 2.5)
    s.set_sdf(sdf)
    s.init_field()
    s.steps(50)
    assert np.abs(s.read_populations()).max() == 0.0
    s3 = O.OracleSolver3D(dict(width=12, height=8, depth=6, precision="fp32", tau=0.55, spongeFraction=0.2,
                               absorb=0.05), 64)
    sdf3 = g.empty_sdf3(12, 8, 6)
    g.add_sphere(sdf3, 12, 8, 6, 5, 3.5, 2.5, 1.8)
    s3.set_sdf(sdf3)
    s3.init_field()
    s3.steps(30)
    assert np.abs(s3.read_populations()).max() == 0.0


def test_zou_he_inlet_velocity_and_outlet_density():
    """step.ts:197-209: Zou-He fixes u = uIn at the inlet column and rho = 1 at the outlet column; the
    TRT collision conserves mass and momentum, so the stored post-collision moments show them."""
    W, H, u = 40, 16, 0.05
    for parabolic in (False, True):
        s = O.OracleSolver(dict(width=W, height=H, left=X_OPEN, right=X_OPEN, bottom=Y_NOSLIP, top=Y_NOSLIP,
                                tau=0.6, uIn=u, parabolic=parabolic, spongeFraction=0.2), 64)
        sdf = g.empty_sdf(W, H)
        g.add_circle(sdf, W, H, 12, 7.5, 2.5)
        s.set_sdf(sdf)
        s.init_field()
        s.steps(30)
        m = s.macro().reshape(H, W, 4)
        ys = np.arange(1, H - 1)
        prof = 4 * u * (ys + 0.5) * (H - ys - 0.5) / H ** 2 if parabolic else np.full(ys.shape, u)
        np.testing.assert_allclose(m[1 : H - 1, 0, 1], prof, atol=1e-13)
        np.testing.assert_allclose(m[1 : H - 1, 0, 2], 0.0, atol=1e-13)
        np.testing.assert_allclose(m[1 : H - 1, W - 1, 0], 1.0, atol=1e-13)


def test_d3q19_inlet_plane_is_equilibrium_and_outlet_rho_one():
    W, H, D = 20, 10, 8
    s = O.OracleSolver3D(dict(width=W, height=H, depth=D, precision="fp32", tau=0.56, uIn=0.07, spongeFraction=0.2,
                              absorb=0.02), 64)
    sdf = g.empty_sdf3(W, H, D)
    g.add_sphere(sdf, W, H, D, 6, 4.5, 3.5, 2)
    s.set_sdf(sdf)
    s.init_field()
    s.steps(25)
    m = s.macro().reshape(D, H, W, 4)
    np.testing.assert_allclose(m[:, :, 0, 0], 0.07, atol=1e-14)
    np.testing.assert_allclose(m[:, :, 0, 3], 1.0, atol=1e-14)
    np.testing.assert_allclose(m[:, :, 0, 1:3], 0.0, atol=1e-14)
    np.testing.assert_allclose(m[:, :, W - 1, 3], 1.0, atol=1e-13)


def test_d3q19_sphere_force_mirror_symmetry():
    """A sphere centred on the channel axis ((H-1)/2, (D-1)/2) with H == D feels equal lateral forces by the
    y <-> z symmetry of the lattice, and its drag is positive."""
    W, H, D = 24, 12, 12
    s = O.OracleSolver3D(dict(width=W, height=H, depth=D, precision="fp32", tau=0.56, uIn=0.05, spongeFraction=0.15,
                              absorb=0.02), 64)
    sdf = g.empty_sdf3(W, H, D)
    g.add_sphere(sdf, W, H, D, 7, (H - 1) / 2, (D - 1) / 2, 2.6)
    s.set_sdf(sdf)
    s.init_field()
    s.steps(60)
    f = s.read_forces()[-1]
    assert f.fx > 0
    assert abs(f.fy) < 1e-10 * abs(f.fx) + 1e-14
    assert abs(f.fz) < 1e-10 * abs(f.fx) + 1e-14
    # all fluid populations finite and the flow symmetric under y -> H-1-y
    m = s.macro().reshape(D, H, W, 4)
    np.testing.assert_allclose(m[:, ::-1, :, 0], m[:, :, :, 0], atol=1e-12)
    np.testing.assert_allclose(m[:, ::-1, :, 1], -m[:, :, :, 1], atol=1e-12)


def test_forcing_ring_cadence_and_reset():
    """forceEvery cadence and the reset semantics of the bookkeeping (src/solver.ts:271-286, 236-241, 315-335)."""
    s = O.OracleSolver(dict(width=16, height=8, left=X_PERIODIC, right=X_PERIODIC, bottom=Y_NOSLIP, top=Y_NOSLIP,
                            tau=0.8, bodyForce=(1e-5, 0), forceEvery=4), 64)
    sdf = g.empty_sdf(16, 8)
    g.add_circle(sdf, 16, 8, 8, 3.5, 1.5)
    s.set_sdf(sdf)
    s.init_field()
    s.steps(4000)
    assert len(s.read_forces()) == 1000
    s.init_field()  # step = 0 and forces reset
    s.steps(40)
    s.steps(400)
    f = s.read_forces()
    assert len(f) == 110 and f[0].step == 4 and f[-1].step == 440
    s.steps(400)
    assert s.read_forces()[0].step == 444
    # ring overrun: more than HISTORY_LEN samples since the last read keeps only the newest 8192
    s.steps(4 * 9000)
    f = s.read_forces()
    assert len(f) == 8192 and f[-1].step == s.step and f[0].step == s.step - 4 * 8191
