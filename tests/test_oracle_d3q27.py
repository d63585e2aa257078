"""D3Q27 velocity set (the reference's roadmap switch, docs/research/3d-lattice-boltzmann-webgpu.md:74)
in the oracle.

The tables are pinned by the properties that define the set: all 27 velocities of {-1, 0, 1}^3
once each, weights summing to 1 with the second- and fourth-order isotropy of the standard D3Q27
quadrature (sum w c_a c_b = delta_ab / 3, sum w c_a c_b c_c c_d = (d_ab d_cd + d_ac d_bd +
d_ad d_bc) / 9, odd moments zero), the D3Q19 set as its first 19 entries in the reference's order
(so odd i and i + 1 stay opposites) and mirrors that negate one component. The solver is pinned
against the naive FP64 implementation (tests/naive_lbm.py), whose D3Q27 set is built
independently in another order, on a tunnel that exercises every boundary rule (inlet and outlet
planes, slip faces with edges and corners, a slip face through a solid, Bouzidi links with
q < 1/2 and q >= 1/2, sponge, absorbing layer, Smagorinsky); and by the uniform-flow fixed point."""
import itertools

import numpy as np
import pytest

from cells_amd import geometry as g
from oracle import oracle as O
from tests import naive_lbm
from tests.parity_cases import case3


def test_d3q27_tables():
    t = O.tables27()
    t19 = O.tables3()
    C = np.stack([t["cx"], t["cy"], t["cz"]], 1)
    assert sorted(map(tuple, C)) == sorted(itertools.product((-1, 0, 1), repeat=3))
    for k in ("cx", "cy", "cz"):
        np.testing.assert_array_equal(t[k][:19], t19[k])
    w = t["w"]
    assert abs(w.sum() - 1) < 1e-15
    eye = np.eye(3)
    for a in range(3):
        assert abs((w * C[:, a]).sum()) < 1e-15
        for b in range(3):
            assert abs((w * C[:, a] * C[:, b]).sum() - eye[a, b] / 3) < 1e-15
            for c in range(3):
                assert abs((w * C[:, a] * C[:, b] * C[:, c]).sum()) < 1e-15
                for d in range(3):
                    iso = (eye[a, b] * eye[c, d] + eye[a, c] * eye[b, d] + eye[a, d] * eye[b, c]) / 9
                    assert abs((w * C[:, a] * C[:, b] * C[:, c] * C[:, d]).sum() - iso) < 1e-15
    for i in range(1, 27):
        assert (C[t["opp"][i]] == -C[i]).all() and t["opp"][i] == (i + 1 if i % 2 else i - 1)
        assert (C[t["mirror_y"][i]] == C[i] * [1, -1, 1]).all()
        assert (C[t["mirror_z"][i]] == C[i] * [1, 1, -1]).all()
    solid, inlet, outlet = O.flag_bits(27)
    assert (solid, inlet, outlet) == (1 << 27, 1 << 28, 1 << 29)


def _naive_order():
    t = O.tables27()
    C = np.stack([t["cx"], t["cy"], t["cz"]], 1)
    return np.array([int(np.where((naive_lbm.C27 == c).all(axis=1))[0][0]) for c in C])


@pytest.mark.parametrize("smag", [0.0, 0.12])
def test_d3q27_oracle_matches_naive(smag):
    cfg, sdf = case3("fp32", W=20, H=12, D=10, smag=smag)
    cfg["velocitySet"] = 27
    # a body on the y = 0, z = D-1 edge: corner directions reflected through both faces meet a solid node
    g.add_sphere(sdf, 20, 12, 10, 11, 0.3, 9.2, 1.5)
    s = O.OracleSolver3D(dict(cfg), 64)
    assert s.Q == 27
    s.set_sdf(sdf)
    s.init_field()
    perm = _naive_order()
    done = 0
    for n in (1, 10, 60):
        s.steps(n - done)
        done = n
        f_naive, forces = naive_lbm.run3d(dict(cfg), sdf, n, naive_lbm.C27, naive_lbm.W27)
        err = np.abs(s.read_populations() - f_naive[perm]).max()
        assert err < 1e-13, (n, err)
    _, vals = s.read_forces_raw()
    assert np.abs(vals - forces).max() < 1e-12 * max(1.0, np.abs(forces).max())
    assert np.abs(forces).max() > 1e-4
    s.destroy()


def test_d3q27_uniform_flow_is_a_fixed_point():
    """No body: the inflow equilibrium everywhere stays put through every boundary rule of the empty
    tunnel (inlet and outlet planes, slip faces, edges, corners, sponge, absorbing layer)."""
    cfg = dict(width=16, height=9, depth=7, precision="fp32", tau=0.53, uIn=0.07, spongeFraction=0.3, spongeTau=1,
               absorb=0.05, velocitySet=27)
    s = O.OracleSolver3D(cfg, 64)
    s.init_field()
    f0 = s.read_populations()
    s.steps(50)
    assert np.abs(s.read_populations() - f0).max() < 1e-15
    m = s.read_macro().reshape(-1, 4)
    np.testing.assert_allclose(m[:, 0], 0.07, rtol=0, atol=1e-15)
    s.destroy()


def test_d3q27_rejects_other_sets():
    cfg, _ = case3("fp32")
    cfg["velocitySet"] = 15
    with pytest.raises(ValueError):
        O.OracleSolver3D(cfg, 64)
