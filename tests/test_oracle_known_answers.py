"""Pins the oracle: the reference's own known-answer unit tests, plus invariants and closed forms.

Known answers are those of src/lattice3d.test.ts:6-76, src/analysis.test.ts:65-72,
src/solver3d.test.ts:121-143 and src/shaders/common3d.test.ts:25-35, asked of the oracle library
itself (its tables, equilibria, parameter derivations and codecs).
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as O

C3 = None


def test_d3q19_weights_and_isotropy():
    """src/lattice3d.test.ts:6-19."""
    t = O.tables3()
    C = [t["cx"], t["cy"], t["cz"]]
    w = t["w"]
    assert len(w) == 19
    assert abs(w.sum() - 1) < 1e-12
    for a in range(3):
        assert abs((w * C[a]).sum()) < 1e-12
        for b in range(3):
            m2 = (w * C[a] * C[b]).sum()
            assert abs(m2 - (1 / 3 if a == b else 0)) < 1e-12


def test_d3q19_opp_and_mirrors():
    """src/lattice3d.test.ts:21-37."""
    t = O.tables3()
    cx, cy, cz, opp, my, mz = (t[k] for k in ("cx", "cy", "cz", "opp", "mirror_y", "mirror_z"))
    for i in range(19):
        assert opp[opp[i]] == i
        assert (cx[opp[i]], cy[opp[i]], cz[opp[i]]) == (-cx[i], -cy[i], -cz[i])
        assert my[my[i]] == i and (cx[my[i]], cy[my[i]], cz[my[i]]) == (cx[i], -cy[i], cz[i])
        assert mz[mz[i]] == i and (cx[mz[i]], cy[mz[i]], cz[mz[i]]) == (cx[i], cy[i], -cz[i])
    # odd i and i + 1 are the TRT pairs and cover 1..18
    seen = set()
    for a in range(1, 19, 2):
        assert opp[a] == a + 1
        seen |= {a, a + 1}
    assert sorted(seen) == list(range(1, 19))


def test_d2q9_tables():
    t = O.tables2()
    cx, cy, w, opp, my = (t[k] for k in ("cx", "cy", "w", "opp", "mirror_y"))
    assert abs(w.sum() - 1) < 1e-15
    assert abs((w * cx * cx).sum() - 1 / 3) < 1e-15 and abs((w * cx * cy).sum()) < 1e-15
    for i in range(9):
        assert opp[opp[i]] == i and (cx[opp[i]], cy[opp[i]]) == (-cx[i], -cy[i])
        assert my[my[i]] == i and (cx[my[i]], cy[my[i]]) == (cx[i], -cy[i])
    # TRT pairs (src/lattice.ts:10-15)
    assert [(a, opp[a]) for a in (1, 2, 5, 6)] == [(1, 3), (2, 4), (5, 7), (6, 8)]


@pytest.mark.parametrize("rho,ux,uy,uz", [(1, 0, 0, 0), (1.02, 0.05, -0.03, 0.01), (0.97, -0.1, 0.02, 0.04)])
def test_equilibrium3_moments(rho, ux, uy, uz):
    """src/lattice3d.test.ts:39-57 (shifted storage: f = feq - w)."""
    t = O.tables3()
    f = np.array([O.feq3(i, rho, ux, uy, uz) for i in range(19)]) + t["w"]
    assert abs(f.sum() - rho) < 1e-12
    for c, u in zip((t["cx"], t["cy"], t["cz"]), (ux, uy, uz)):
        assert abs((f * c).sum() - rho * u) < 1e-12


def test_equilibrium2_moments():
    """src/analysis.test.ts:65-72."""
    t = O.tables2()
    f = np.array([O.feq2(i, 1.1, 0.05, -0.02) for i in range(9)]) + t["w"]
    assert abs(f.sum() - 1.1) < 1e-12
    assert abs((f * t["cx"]).sum() - 1.1 * 0.05) < 1e-12
    assert abs((f * t["cy"]).sum() - 1.1 * -0.02) < 1e-12
    # second moment: sum f c_x c_x = rho (1/3 + ux^2)
    assert abs((f * t["cx"] ** 2).sum() - 1.1 * (1 / 3 + 0.05 ** 2)) < 1e-12


def test_flag_bits_disjoint():
    """src/lattice3d.test.ts:59-70."""
    dirs = ((1 << 19) - 1) & ~1
    s, i, o = O.flag_bits(3)
    for f in (s, i, o):
        assert f & dirs == 0 and 0 < f < 2 ** 32
    assert s & i == 0 and s & o == 0 and i & o == 0
    s2, i2, o2 = O.flag_bits(2)
    assert (s2, i2, o2) == (1 << 9, 1 << 10, 1 << 11)


def test_trt_lambda3_known_answers():
    """src/lattice3d.test.ts:72-76."""
    assert O.trt_lambda3(0.6) == 3 / 16
    assert abs(O.trt_lambda3(0.5288) - 50 * 0.0288 ** 2) < 1e-12
    assert abs(O.trt_lambda3(0.51) - 0.005) < 1e-12


def test_params3_derivation():
    """src/solver3d.test.ts:121-143: spongeStart = (W-1)(1-0.15), uIn, tau, smagC2 = Cs^2."""
    s = O.OracleSolver3D(dict(width=8, height=6, depth=4, precision="fp32", tau=0.6, uIn=0, spongeFraction=0.15), 32)
    p = s.params()
    assert abs(p[0] - 0.6) < 1e-6 and abs(p[4] - 7 * 0.85) < 1e-5
    assert abs(p[1] - 3 / 16) < 1e-7  # lambda follows trtLambda3(0.6)
    assert p[5] == 1.0  # tauSponge = max(spongeTau, tau)
    s.set_inlet(0.05)
    assert abs(s.params()[3] - 0.05) < 1e-6
    s.set_tau(0.7)
    assert abs(s.params()[0] - 0.7) < 1e-6
    s.set_smagorinsky(0.16)
    assert abs(s.params()[2] - 0.0256) < 1e-6
    # no sponge: spongeStart = 1e9 (src/solver3d.ts:204)
    s2 = O.OracleSolver3D(dict(width=8, height=6, depth=4, precision="fp32", tau=0.52), 32)
    assert s2.params()[4] == 1e9 and abs(s2.params()[1] - 50 * 0.02 ** 2) < 1e-8


def test_params2_derivation():
    """src/solver.ts:180-197."""
    s = O.OracleSolver(dict(width=512, height=128, left=2, right=2, bottom=2, top=2, tau=0.548, spongeFraction=0.15,
                            absorb=0.02, inletLayerFraction=0.08, bodyForce=(1e-6, 0)), 64)
    p = s.params()
    assert abs(p[4] - 511 * 0.85) < 1e-9  # spongeStart
    assert p[5] == 1.0  # tauSponge = max(1, 0.548)
    assert abs(p[10] - 0.08 * 511) < 1e-9  # inletLayer
    assert p[12] == 1 and p[7] == 1e-6  # forcing on
    assert abs(p[1] - 3 / 16) < 1e-15


def test_f16_codec_matches_numpy_rne():
    """The oracle's float32 -> float16 conversion is IEEE round-to-nearest-even, subnormals included."""
    rng = np.random.default_rng(0)
    xs = np.concatenate([
        rng.standard_normal(20000).astype(np.float32) * 1000,
        (rng.standard_normal(20000) * 1e-5).astype(np.float32),
        (rng.standard_normal(2000) * 1e-7).astype(np.float32),
        np.array([0.0, -0.0, 65504, 65519.99, 65520, -65520, 6.1e-5, 5.96e-8, 2.98e-8, 2.99e-8, 1e-9, np.inf, -np.inf],
                 dtype=np.float32),
    ])
    ref = xs.astype(np.float16).view(np.uint16)
    got = np.array([O.f32_to_f16(float(x)) for x in xs], dtype=np.uint16)
    np.testing.assert_array_equal(got, ref)


def test_bf16_codec_matches_torch_rne():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 10,
                         (rng.standard_normal(5000) * 1e-3).astype(np.float32),
                         np.array([0.0, -0.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 3.0e38, -1e-40], dtype=np.float32)])
    ref = torch.from_numpy(xs).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    got = np.array([O.f32_to_bf16(float(x)) for x in xs], dtype=np.uint16)
    np.testing.assert_array_equal(got, ref)


def test_fp16_storage_roundtrip_error_bound():
    """common3d.ts:95,100 codec: |decode(encode(v)) - v| <= 1/2 ulp_f16(v * 2^15) / 2^15 inside +-1.99."""
    rng = np.random.default_rng(2)
    v = (rng.standard_normal(5000) * 0.05).astype(np.float32)
    enc = np.array([O.f32_to_f16(float(np.clip(x, -1.99, 1.99) * 32768.0)) for x in v], dtype=np.uint16)
    dec = enc.view(np.float16).astype(np.float32) / np.float32(32768.0)
    scaled = np.abs(v.astype(np.float64) * 32768.0)
    ulp = np.where(scaled < 2 ** -14, 2.0 ** -24, 2.0 ** (np.floor(np.log2(np.maximum(scaled, 2 ** -14))) - 10))
    assert np.all(np.abs(dec - v) <= 0.5 * ulp / 32768.0 + 1e-12)


def test_f16_stochastic_rounding_closed_forms():
    """The stochastic conversion (LBM_FP16_SR) against what rounding with a uniform 13-bit dither fixes:
    every result is one of the two binary16 neighbours of x (RNE's candidates), a representable x is
    returned for every word, the mean over all 8192 dither values is x exactly (no bias), the result
    is odd in x, and below 2^-14 it is round-to-nearest-even."""
    rng = np.random.default_rng(5)
    xs = np.concatenate([rng.standard_normal(40).astype(np.float32) * 1000,
                         (rng.standard_normal(40) * 0.01).astype(np.float32),
                         np.array([6.103515625e-05, 65216.0, 1.0 + 2 ** -12, 2048.0 - 2 ** -3], dtype=np.float32)])
    for x in xs:
        if abs(x) < 2 ** -14:
            continue
        lo = np.float16(x)  # RNE result: one of the neighbours
        outs = np.array([O.f32_to_f16_sr(float(x), int(d) << 19) for d in range(8192)], dtype=np.uint16)
        vals = outs.view(np.float16).astype(np.float64)
        assert set(np.unique(vals)) <= {float(lo), float(np.nextafter(lo, np.float16(np.inf))),
                                        float(np.nextafter(lo, np.float16(-np.inf)))}
        assert abs(vals.mean() - np.float64(x)) <= 1e-12 * abs(np.float64(x)), (x, vals.mean())
        assert O.f32_to_f16_sr(float(-x), 12345 << 19) == O.f32_to_f16_sr(float(x), 12345 << 19) ^ 0x8000
    for x in np.array([1.0, 0.5, 1024.0, -3.25, 65504.0 / 2], dtype=np.float32):
        assert {O.f32_to_f16_sr(float(x), int(r)) for r in rng.integers(0, 2 ** 32, 64)} == {O.f32_to_f16(float(x))}
    for x in (1e-6, -3e-7, 5.9e-8, 0.0):
        for r in (0, 2 ** 31, 2 ** 32 - 1):
            assert O.f32_to_f16_sr(x, r) == O.f32_to_f16(x)
    # the random words of distinct (node, step, population) differ and use all of their top 13 bits
    words = np.array([O.step_random(n, s, i) for n in range(40) for s in range(5) for i in range(19)], dtype=np.uint64)
    assert len(np.unique(words)) == words.size
    top = words >> np.uint64(19)
    assert abs(top.mean() / 4095.5 - 1) < 0.05 and top.min() < 400 and top.max() > 7800
