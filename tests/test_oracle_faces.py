"""Zou-He faces of Hecht and Harting (arXiv:0811.4593) on the D3Q19 x planes, in the FP64 oracle.

The rebuilt populations are fixed by the moments they must give: every inlet node (x = 0) carries
exactly u = (uIn, 0, 0) with its own density, every outlet node (x = W - 1) exactly rho = 1 with no
transverse velocity. Where the collision conserves mass and momentum (no absorbing source there,
absorb = 0 at the outlet), the stored post-collision populations keep those moments, so readMacro
shows them to FP64 rounding. A sign slip in the transverse correction shows as u_y, u_z != 0, a
wrong density closure as u_x != uIn, and bodies next to both planes make the transverse flow there
non-zero. The reference's equilibrium planes give the same moments, so the test also checks what
tells the faces apart: the inlet density floats and the face populations are not at equilibrium."""
import numpy as np

from cells_amd import geometry as g
from oracle.oracle import OracleSolver3D, feq3


def _tunnel(faces, absorb=0.0):
    W, H, D, u = 22, 11, 9, 0.05
    cfg = dict(width=W, height=H, depth=D, precision="fp32", tau=0.56, uIn=u, spongeFraction=0.15, spongeTau=1,
               absorb=absorb)
    sdf = g.empty_sdf3(W, H, D)
    g.add_sphere(sdf, W, H, D, 3.3, 4.6, 3.9, 2.1)  # two cells from the inlet plane
    g.add_sphere(sdf, W, H, D, W - 4.2, 6.1, 4.8, 2.0)  # two cells from the outlet plane
    s = OracleSolver3D(cfg, 64)
    s.set_sdf(sdf)
    s.init_field()
    s.set_open_faces(faces)
    return s, W, H, D, u


def _planes(s, W, H, D):
    m = s.read_macro().reshape(D, H, W, 4)  # (ux, uy, uz, rho)
    return m[:, :, 0], m[:, :, W - 1], m


def test_zou_he_faces_carry_prescribed_moments():
    s, W, H, D, u = _tunnel("zou-he")
    for n in (1, 30, 120):
        s.steps(n if n == 1 else n - s.step)
        inlet, outlet, m = _planes(s, W, H, D)
        solid = m[..., 3] == 0
        assert not solid[:, :, 0].any() and not solid[:, :, W - 1].any()
        np.testing.assert_allclose(inlet[..., 0], u, rtol=0, atol=1e-13)
        np.testing.assert_allclose(inlet[..., 1:3], 0, rtol=0, atol=1e-13)
        np.testing.assert_allclose(outlet[..., 3], 1, rtol=0, atol=1e-13)
        np.testing.assert_allclose(outlet[..., 1:3], 0, rtol=0, atol=1e-13)
    # the bodies drive transverse flow next to both planes, so the corrections were exercised
    assert np.abs(m[:, :, 1, 1:3]).max() > 1e-4 and np.abs(m[:, :, W - 2, 1:3]).max() > 1e-4
    # not the equilibrium planes: the inlet density floats, the outlet velocity is the node's own
    assert np.abs(inlet[..., 3] - 1).max() > 1e-5
    f = s.read_populations().reshape(19, D, H, W)
    y, z = H // 2, D // 2
    eq = np.array([feq3(i, inlet[z, y, 3], u, 0.0, 0.0) for i in range(19)])
    assert np.abs(f[:, z, y, 0] - eq).max() > 1e-6
    s.destroy()


def test_equilibrium_planes_unchanged_by_the_switch():
    a, W, H, D, _ = _tunnel("equilibrium", absorb=0.02)
    b, *_ = _tunnel("zou-he", absorb=0.02)
    b.set_open_faces("equilibrium")
    a.steps(40)
    b.steps(40)
    np.testing.assert_array_equal(a.read_populations(), b.read_populations())
    inlet, outlet, _ = _planes(a, W, H, D)
    np.testing.assert_allclose(inlet[..., 3], 1, rtol=0, atol=1e-13)
    a.destroy()
    b.destroy()


def test_zou_he_faces_settle_to_the_same_flow():
    """Away from the planes, a steady flow past the bodies differs little between the two open
    boundaries (both impose the same uniform inflow and rho = 1 outflow)."""
    sims = {k: _tunnel(k, absorb=0.02)[0] for k in ("equilibrium", "zou-he")}
    for s in sims.values():
        s.steps(600)
    W = 22
    ma, mb = (s.read_macro().reshape(9, 11, W, 4) for s in sims.values())
    mid = slice(6, W - 6)
    assert np.isfinite(mb).all()
    assert np.abs(ma[:, :, mid, 0] - mb[:, :, mid, 0]).max() < 0.1 * 0.05
    for s in sims.values():
        s.destroy()
