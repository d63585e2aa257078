"""Host-side diagnostics of the finite-difference checks of the adjoint: the Bouzidi links of a geometry
and the links that change interpolation branch between two geometries (cells_amd.adjoint)."""
import numpy as np

from cells_amd import shapes
from cells_amd.adjoint import bouzidi_branch_switches, bouzidi_links


def _one_solid_2d(v_solid, v_fluid=0.5, W=9, H=7):
    sdf = np.full((H, W), v_fluid, dtype=np.float32)
    sdf[3, 4] = v_solid
    return sdf.reshape(-1)


def test_links_of_one_solid_node_2d():
    """One solid node in fluid: each of its 8 neighbours has one link, pulling from it; with the SDF
    -1/2 there and +1/2 at the neighbours, q = 0.5 / (0.5 + 0.5) = 1/2 exactly."""
    W, H = 9, 7
    L = bouzidi_links(_one_solid_2d(-0.5), W, H)
    nodes = sorted(int(n) for i in L for n in L[i][0])
    assert len(nodes) == 8
    expect = sorted((4 + dx) + W * (3 + dy) for dx in (-1, 0, 1) for dy in (-1, 0, 1) if dx or dy)
    assert nodes == expect
    for i, (idx, q) in L.items():
        if len(idx):
            assert len(idx) == 1
            assert q[0] == np.float32(0.5)
            # the link's node is the solid node plus c_i (the node pulls from x - c_i)
            cx, cy = [(0, 0), (1, 0), (0, 1), (-1, 0), (0, -1), (1, 1), (-1, 1), (-1, -1), (1, -1)][i]
            assert idx[0] == (4 + cx) + W * (3 + cy)


def test_branch_switches_count_q_crossing_half_clamps_and_flips():
    W, H = 9, 7
    a = _one_solid_2d(-0.5)
    assert bouzidi_branch_switches(a, a, W, H) == 0
    # q = 0.5 / 1.1 < 1/2 on all eight links: every link changes branch
    assert bouzidi_branch_switches(a, _one_solid_2d(-0.6), W, H) == 8
    # q = 0.5 / 0.9 > 1/2: the same branch as q = 1/2
    assert bouzidi_branch_switches(a, _one_solid_2d(-0.4), W, H) == 0
    # a node almost on the surface: q = 2e-6 clamped to 1e-3, against q = 0.0196 unclamped
    assert bouzidi_branch_switches(_one_solid_2d(-0.5, 1e-6), _one_solid_2d(-0.5, 1e-2), W, H) == 8
    # a second solid node next to the first: links appear and disappear (a flip)
    b = a.reshape(H, W).copy()
    b[3, 5] = -0.5
    assert bouzidi_branch_switches(a, b.reshape(-1), W, H) > 0
    # the open x planes are never solid
    c = np.full((H, W), 0.5, dtype=np.float32)
    c[:, 0] = -1.0
    assert sum(len(v[0]) for v in bouzidi_links(c.reshape(-1), W, H).values()) == 0


def test_links_3d_sphere_counts_and_symmetry():
    """D3Q19 links of a sphere centred between lattice nodes in y and z: the link set is symmetric under
    y -> H - 1 - y, and q in [1e-3, 1]."""
    W, H, D = 24, 16, 16
    sdf, _ = shapes.sphere_sh(W, H, D, (10.0, 7.5, 7.5), 4.0, np.zeros(16))
    L = bouzidi_links(sdf, W, H, D)
    total = sum(len(v[0]) for v in L.values())
    assert total > 0
    mirror = {1: 1, 2: 2, 3: 4, 4: 3, 5: 5, 6: 6, 7: 13, 8: 14, 9: 9, 10: 10, 11: 18, 12: 17, 13: 7, 14: 8,
              15: 15, 16: 16, 17: 12, 18: 11}
    for i, (idx, q) in L.items():
        assert np.all((q >= 1e-3) & (q <= 1.0))
        x, y, z = idx % W, (idx // W) % H, idx // (W * H)
        m_idx, m_q = L[mirror[i]]
        got = dict(zip(m_idx.tolist(), m_q.tolist()))
        for xx, yy, zz, qq in zip(x, y, z, q):
            j = int(xx + W * ((H - 1 - yy) + H * zz))
            assert j in got and got[j] == qq
    assert bouzidi_branch_switches(sdf, sdf, W, H, D) == 0


def test_naca_case_switches_only_on_a1_at_h_1e3():
    """The 2D shape-gradient case (384 x 128 NACA0012 with thickness bumps): between theta +- 1e-3 e_p
    exactly one Bouzidi link changes branch, for the first bump amplitude a_1, and none at h = 1e-4;
    this is why the h = 1e-3 central difference of a_1 is not J's derivative there."""
    W, H, chord = 384, 128, 64.0
    theta = np.array([5.0] + [0.0] * 8)

    def sdf(th):
        return shapes.naca_bumps(W, H, 96.0, 63.5, chord, th[0], th[1:])[0]

    for h, expect in ((1e-3, [0, 1, 0, 0, 0, 0, 0, 0, 0]), (1e-4, [0] * 9)):
        got = []
        for p in range(9):
            e = np.zeros(9)
            e[p] = h
            got.append(bouzidi_branch_switches(sdf(theta + e), sdf(theta - e), W, H))
        assert got == expect, (h, got)
