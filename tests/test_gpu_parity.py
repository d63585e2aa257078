"""GPU (gfx950 HIP library, through the C ABI) against the oracle, on the same inputs.

Bit-exact where the work is integer (flag words, force slots, node classification); within
stated tolerances for FP32 arithmetic, whose operation order the GPU compiler may contract into
FMAs. Each float comparison also prints the error of both the GPU and the FP32 oracle against the
FP64 oracle, and requires the GPU's to be of the same size.
"""
from __future__ import annotations

import numpy as np
import pytest

from cells_amd import geometry as g
from cells_amd.lbm import Solver, Solver3D
from oracle.oracle import OracleSolver, OracleSolver3D, flag_bits, tables3
from tests.parity_cases import CASES2, case3, taylor_green_macro

pytestmark = pytest.mark.gpu

# Max |f_gpu - f_oracle32| per step count, FP32 storage. Populations are stored as f - w, of size
# ~1e-2 here; these bounds are ~100-1000 float32 ulps at that size after 100 steps.
TOL32 = {1: 2e-6, 10: 5e-6, 100: 3e-5}
# FP16 x 2^15 storage: one quantum at |f - w| ~ 0.03 is 2^-11 / 2^15 * 2^... ~ 1.5e-5; allow a few.
TOL16 = {1: 8e-5, 10: 2e-4, 100: 6e-4}
# BF16: 8 significant bits, one quantum at 0.03 is ~1.2e-4.
TOLBF = {1: 5e-4, 10: 1.5e-3, 100: 4e-3}


def _build2(name, real=None):
    cfg, sdf_fn = CASES2[name]
    W, H = cfg["width"], cfg["height"]
    sdf = sdf_fn(W, H) if sdf_fn else None
    sims = {}
    sims["gpu"] = Solver.create(dict(cfg))
    sims["o32"] = OracleSolver(dict(cfg), 32)
    sims["o64"] = OracleSolver(dict(cfg), 64)
    for s in sims.values():
        if sdf is not None:
            s.set_sdf(sdf)
        if name == "box":
            s.init_field(taylor_green_macro(W))
        else:
            s.init_field()
    return sims


def _err(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64))))


@pytest.mark.parametrize("name", list(CASES2))
def test_d2q9_flags_bit_exact(name):
    sims = _build2(name)
    fg = sims["gpu"].read_flags()
    fo = sims["o32"].read_flags()
    assert fg.dtype == fo.dtype == np.uint32
    np.testing.assert_array_equal(fg, fo)


@pytest.mark.parametrize("name", list(CASES2))
def test_d2q9_fields_match_oracle(name):
    sims = _build2(name)
    done = 0
    for n in (1, 10, 100):
        for s in sims.values():
            s.run(n - done) if hasattr(s, "sync") else s.steps(n - done)
        done = n
        fg = sims["gpu"].read_populations()
        f32 = sims["o32"].read_populations()
        f64 = sims["o64"].read_populations()
        e = _err(fg, f32)
        eg64, eo64 = _err(fg, f64), _err(f32, f64)
        print(f"{name} step {n}: |gpu-o32| {e:.3e}  |gpu-o64| {eg64:.3e}  |o32-o64| {eo64:.3e}")
        assert np.isfinite(fg).all()
        assert e <= TOL32[n], (name, n, e)
        assert eg64 <= 3 * eo64 + TOL32[n], (name, n, eg64, eo64)


@pytest.mark.parametrize("name", ["tunnel", "channel", "periodic"])
def test_d2q9_forces_match_oracle(name):
    sims = _build2(name)
    for s in sims.values():
        s.run(300) if hasattr(s, "sync") else s.steps(300)
    sg, vg = sims["gpu"].read_forces_raw()
    so, vo = sims["o32"].read_forces_raw()
    s6, v6 = sims["o64"].read_forces_raw()
    np.testing.assert_array_equal(sg, so)
    assert len(sg) == 300 and sg[0] == 1 and sg[-1] == 300
    scale = np.abs(v6).max()
    e = np.abs(vg - vo).max() / scale
    print(f"{name} force: rel |gpu-o32| {e:.3e}, rel |gpu-o64| {np.abs(vg - v6).max() / scale:.3e}, "
          f"rel |o32-o64| {np.abs(vo - v6).max() / scale:.3e}")
    assert e < 1e-4


def test_d2q9_macro_matches_oracle():
    sims = _build2("tunnel")
    sims["gpu"].run(50, 20)
    sims["o32"].steps(50)
    mg = sims["gpu"].read_macro().reshape(-1, 4)
    mo = sims["o32"].macro().reshape(-1, 4)
    np.testing.assert_array_equal(mg[:, 3], mo[:, 3])  # solid marker
    assert _err(mg, mo) < 2e-5


def test_d2q9_geometry_change_refills():
    """setSdf after flow has developed: cells the body uncovered are refilled from their neighbours."""
    name = "tunnel"
    sims = _build2(name)
    for s in sims.values():
        s.run(40) if hasattr(s, "sync") else s.steps(40)
    cfg = CASES2[name][0]
    W, H = cfg["width"], cfg["height"]
    sdf2 = g.empty_sdf(W, H)
    g.add_circle(sdf2, W, H, 20, 11.5, 3)
    for s in sims.values():
        s.set_sdf(sdf2)
    np.testing.assert_array_equal(sims["gpu"].read_flags(), sims["o32"].read_flags())
    e = _err(sims["gpu"].read_populations(), sims["o32"].read_populations())
    assert e < 1e-5
    for s in sims.values():
        s.run(20) if hasattr(s, "sync") else s.steps(20)
    assert _err(sims["gpu"].read_populations(), sims["o32"].read_populations()) < 2e-5


@pytest.mark.parametrize("precision", ["fp32", "fp16", "bf16", "fp16sr"])
def test_d3q19_fields_match_oracle(precision):
    cfg, sdf = case3(precision)
    gpu = Solver3D.create(dict(cfg))
    o32 = OracleSolver3D(dict(cfg), 32)
    o64 = OracleSolver3D(dict(cfg), 64)
    for s in (gpu, o32, o64):
        s.set_sdf(sdf)
        s.init_field()
    np.testing.assert_array_equal(gpu.read_flags(), o32.read_flags())
    tol = {"fp32": TOL32, "fp16": TOL16, "bf16": TOLBF, "fp16sr": TOL16}[precision]
    done = 0
    for n in (1, 10, 100):
        gpu.run(n - done)
        o32.steps(n - done)
        o64.steps(n - done)
        done = n
        fg = gpu.read_populations()
        fo = o32.read_populations()
        f6 = o64.read_populations()
        e = _err(fg, fo)
        print(f"d3q19 {precision} step {n}: |gpu-o32| {e:.3e} |gpu-o64| {_err(fg, f6):.3e} |o32-o64| {_err(fo, f6):.3e}")
        assert np.isfinite(fg).all()
        assert e <= tol[n], (precision, n, e)
    sg, vg = gpu.read_forces_raw()
    so, vo = o32.read_forces_raw()
    np.testing.assert_array_equal(sg, so)
    rel = np.abs(vg - vo).max() / np.abs(vo).max()
    print(f"d3q19 {precision} force rel err {rel:.3e}")
    assert rel < {"fp32": 1e-4, "fp16": 2e-3, "bf16": 2e-2, "fp16sr": 2e-3}[precision]


@pytest.mark.parametrize("precision", ["fp32", "fp16", "bf16"])
def test_d3q19_macro_matches_oracle(precision):
    """readMacro (src/solver3d.ts:290-292, src/shaders/aux3d.ts:128-152): (ux, uy, uz, rho) per cell,
    all four zero in solids, from the same populations as the oracle's."""
    cfg, sdf = case3(precision)
    gpu = Solver3D.create(dict(cfg))
    o32 = OracleSolver3D(dict(cfg), 32)
    for s in (gpu, o32):
        s.set_sdf(sdf)
        s.init_field()
    gpu.run(60, 25)
    o32.steps(60)
    m = gpu.read_macro().reshape(-1, 4)
    mo = o32.read_macro().reshape(-1, 4)
    solid = (o32.read_flags() & np.uint32(flag_bits(19)[0])) != 0
    assert solid.sum() > 50
    np.testing.assert_array_equal(m[solid], 0.0)
    assert np.all(m[~solid, 3] > 0.9)
    # the macro of the GPU's own populations, decoded, in float64: the layout is (u, rho)
    f = gpu.read_populations().astype(np.float64)
    t = tables3()
    rho = 1.0 + f.sum(0)
    own = np.stack([(t["cx"] @ f) / rho, (t["cy"] @ f) / rho, (t["cz"] @ f) / rho, rho], 1)
    assert np.abs(m[~solid] - own[~solid]).max() < 2e-6
    e = np.abs(m - mo).max()
    print(f"d3q19 {precision} macro |gpu-o32| {e:.3e}")
    assert e < {"fp32": 1e-6, "fp16": 1e-3, "bf16": 5e-3}[precision]
    gpu.destroy()
    o32.destroy()


@pytest.mark.parametrize("q,precision", [(19, "fp32"), (19, "fp16"), (27, "fp32"), (27, "fp16")])
def test_d3_geometry_change_refills(q, precision):
    """setSdf on a developed 3D flow (src/solver3d.ts:211-228, src/shaders/aux3d.ts:83-99): the body
    moves downstream and a cavity opens inside the old one, so refilled cells take either the mean of
    their formerly fluid neighbours or the inflow equilibrium. Flags bit-exact with the oracle, the
    refilled populations and 20 more steps within the parity bounds."""
    cfg, sdf = case3(precision)
    cfg["velocitySet"] = q
    W, H, D = cfg["width"], cfg["height"], cfg["depth"]
    gpu = Solver3D.create(dict(cfg))
    o32 = OracleSolver3D(dict(cfg), 32)
    for s in (gpu, o32):
        s.set_sdf(sdf)
        s.init_field()
    gpu.run(40)
    o32.steps(40)
    solid_bit = np.uint32(flag_bits(q)[0])
    solid0 = (o32.read_flags() & solid_bit) != 0
    sdf2 = g.empty_sdf3(W, H, D)
    g.add_sphere(sdf2, W, H, D, 10.0, 6.5, 5.5, 3.2)
    z, y, x = np.mgrid[0:D, 0:H, 0:W].astype(np.float64)
    cavity = np.sqrt((x - 8.0) ** 2 + (y - 6.5) ** 2 + (z - 5.5) ** 2) - 0.8
    sdf2 = np.maximum(sdf2.reshape(D, H, W), -cavity).astype(np.float32).reshape(-1)
    g.add_sphere(sdf2, W, H, D, 17, 4, 0.5, 2.2)
    for s in (gpu, o32):
        s.set_sdf(sdf2)
    fg = gpu.read_flags()
    np.testing.assert_array_equal(fg, o32.read_flags())
    uncovered = solid0 & ((fg & solid_bit) == 0)
    assert uncovered.sum() >= 10
    tol = {"fp32": TOL32, "fp16": TOL16}[precision]
    e = _err(gpu.read_populations(), o32.read_populations())
    er = _err(gpu.read_populations()[:, uncovered], o32.read_populations()[:, uncovered])
    print(f"d3q{q} {precision} refill: |gpu-o32| {e:.3e}, on the {int(uncovered.sum())} refilled cells {er:.3e}")
    assert e <= tol[100] and er <= tol[1]
    gpu.run(20)
    o32.steps(20)
    e = _err(gpu.read_populations(), o32.read_populations())
    print(f"d3q{q} {precision} 20 steps after the refill: |gpu-o32| {e:.3e}")
    assert np.isfinite(gpu.read_populations()).all()
    assert e <= tol[100]
    sg, vg = gpu.read_forces_raw()
    so, vo = o32.read_forces_raw()
    np.testing.assert_array_equal(sg, so)
    assert np.abs(vg - vo).max() / np.abs(vo).max() < {"fp32": 1e-4, "fp16": 2e-3}[precision]
    gpu.destroy()
    o32.destroy()


def test_d3q19_smagorinsky_and_setters():
    cfg, sdf = case3("fp32", smag=0.12)
    gpu = Solver3D.create(dict(cfg))
    o32 = OracleSolver3D(dict(cfg), 32)
    for s in (gpu, o32):
        s.set_sdf(sdf)
        s.init_field()
    gpu.run(20)
    o32.steps(20)
    for s in (gpu, o32):
        s.set_tau(0.56)
        s.set_inlet(0.08)
        s.set_smagorinsky(0.05)
    gpu.run(20)
    o32.steps(20)
    assert _err(gpu.read_populations(), o32.read_populations()) < 2e-5


@pytest.mark.parametrize("precision", ["fp16", "bf16"])
def test_d3q19_fp16_bits_match_oracle_codec(precision):
    """After init (no arithmetic differences yet) the stored FP16 / BF16 words equal the oracle's
    codec (round to nearest even; a truncating or biased encoder would miss on about half)."""
    cfg, sdf = case3(precision)
    gpu = Solver3D.create(dict(cfg))
    o32 = OracleSolver3D(dict(cfg), 32)
    for s in (gpu, o32):
        s.set_sdf(sdf)
        s.init_field()
    raw_g = gpu.read_populations_raw()
    raw_o = o32.read_raw().view(np.uint16)
    mism = np.count_nonzero(raw_g != raw_o)
    assert mism <= raw_g.size // 1000, mism


def test_d3q19_fp16sr_step_bits_match_oracle():
    """One step in LBM_FP16_SR: the stored words equal the oracle's (same random words, same stochastic
    rounding) except where the GPU's contracted arithmetic moved a value across a rounding decision,
    and they differ from round-to-nearest on a large share of the values (the rounding is live)."""
    cfg, sdf = case3("fp16sr")
    runs = {}
    for name, sim in (("gpu", Solver3D.create(dict(cfg))), ("o32", OracleSolver3D(dict(cfg), 32)),
                      ("rne", Solver3D.create(dict(cfg, precision="fp16")))):
        sim.set_sdf(sdf)
        sim.init_field()
        if name == "o32":
            sim.steps(1)
            runs[name] = sim.read_raw().view(np.uint16)
        else:
            sim.run(1)
            runs[name] = sim.read_populations_raw()
        sim.destroy()
    diff_oracle = np.count_nonzero(runs["gpu"] != runs["o32"]) / runs["gpu"].size
    diff_rne = np.count_nonzero(runs["gpu"] != runs["rne"]) / runs["gpu"].size
    print(f"fp16sr step words: differ from the oracle {diff_oracle:.2e}, from round-to-nearest {diff_rne:.3f}")
    assert diff_oracle < 5e-3
    assert diff_rne > 0.1


def test_ring_and_reset_semantics_on_gpu():
    """readForces across resetForces (src/solver.ts:315-335) with the GPU ring."""
    cfg, sdf_fn = CASES2["tunnel"]
    s = Solver.create(dict(cfg, forceEvery=4))
    s.set_sdf(sdf_fn(cfg["width"], cfg["height"]))
    s.init_field()
    s.run(40)
    f = s.read_forces()
    assert [x.step for x in f] == list(range(4, 41, 4))
    assert s.read_forces() == []
    s.run(8)
    s.reset_forces()
    assert s.read_forces() == []
    s.run(9)
    f = s.read_forces()
    assert [x.step for x in f] == [52, 56]


def test_d2q9_brush_strokes_match_oracle():
    """brush / brushStroke (src/solver.ts:243-268) during developed flow: a stroke that grows the body
    and an erase that uncovers part of it. Flags bit-exact with the oracle, refilled populations and
    20 more steps within the setSdf bounds, and the brushed field equal bit for bit to the same SDF
    given through setSdf."""
    name = "tunnel"
    sims = _build2(name)
    for s in sims.values():
        s.run(40) if hasattr(s, "sync") else s.steps(40)
    stroke = [(22.0, 8.0), (23.5, 9.0), (25.0, 10.0), (26.5, 11.0)]
    for s in sims.values():
        s.brush_stroke(stroke, 2.0)
        s.brush(12.5, 11.5, 1.8, erase=True)
    np.testing.assert_array_equal(sims["gpu"].read_flags(), sims["o32"].read_flags())
    assert _err(sims["gpu"].read_populations(), sims["o32"].read_populations()) < 1e-5
    for s in sims.values():
        s.run(20) if hasattr(s, "sync") else s.steps(20)
    assert _err(sims["gpu"].read_populations(), sims["o32"].read_populations()) < 2e-5
    # the same field through setSdf on a second GPU solver that took the same path
    cfg, sdf_fn = CASES2[name]
    W, H = cfg["width"], cfg["height"]
    twin = Solver.create(dict(cfg))
    twin.set_sdf(sdf_fn(W, H))
    twin.init_field()
    twin.run(40)
    brushed = sims["o32"].read_sdf()
    twin.set_sdf(brushed)
    twin.run(20)
    np.testing.assert_array_equal(twin.read_flags(), sims["gpu"].read_flags())
    np.testing.assert_array_equal(twin.read_populations(), sims["gpu"].read_populations())
    # an empty box is a no-op; non-finite input and 3D solvers are refused
    f0 = sims["gpu"].read_flags()
    sims["gpu"].brush(-50.0, 5.0, 2.0)
    np.testing.assert_array_equal(sims["gpu"].read_flags(), f0)
    with pytest.raises(RuntimeError):
        sims["gpu"].brush(float("nan"), 5.0, 2.0)
    s3 = Solver3D.create(dict(width=16, height=8, depth=8, precision="fp32", tau=0.6))
    with pytest.raises(RuntimeError):
        Solver.brush(s3, 4.0, 4.0, 1.0)  # the brush is 2D only
    for s in (twin, s3, sims["gpu"]):
        s.destroy()


@pytest.mark.parametrize("name", ["tunnel", "channel", "periodic"])
def test_d2q9_launch_shapes_give_identical_bits(name):
    """The D2Q9 launch shapes the autotuner picks from (block widths that split the rows differently,
    rows walked bottom-up or top-down on alternate steps): cells are independent within a step and the
    force sample is summed per wave of body nodes in node order, so populations and force samples are
    the same bits in every shape; the autotuner restores the state it times on."""
    cfg, sdf_fn = CASES2[name]
    cfg = dict(cfg, width=150, height=40)
    W, H = cfg["width"], cfg["height"]
    sdf = sdf_fn(W, H)
    outs = {}
    for block, variant in ((64, 0), (128, 1), (256, 2), (192, 1), (64, 2), ("auto", None)):
        s = Solver.create(dict(cfg))
        s.set_sdf(sdf)
        s.init_field()
        early = (np.zeros(0, np.int64), np.zeros((0, 2)))
        if block == "auto":
            s.run(7)
            f7 = s.read_populations()
            early = s.read_forces_raw()
            s.autotune(5)
            np.testing.assert_array_equal(s.read_populations().view(np.uint32), f7.view(np.uint32))
            assert len(s.read_forces_raw()[0]) == 0
            assert s.launch_info()[1] in (0, 1, 2)
            s.run(33)
        else:
            s.set_launch(block, variant)
            assert s.launch_info()[1] == variant
            s.run(40)
        late = s.read_forces_raw()
        outs[(block, variant)] = (s.read_populations(), np.concatenate([early[0], late[0]]),
                                  np.concatenate([early[1], late[1]]))
        s.destroy()
    ref = outs[(64, 0)]
    for k, v in outs.items():
        np.testing.assert_array_equal(v[0].view(np.uint32), ref[0].view(np.uint32), err_msg=str(k))
        np.testing.assert_array_equal(v[1], ref[1], err_msg=str(k))
        np.testing.assert_array_equal(v[2].view(np.uint64), ref[2].view(np.uint64), err_msg=str(k))


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_d3q19_launch_shapes_give_identical_bits(precision):
    """The launch shapes the autotuner picks from (a row of 150 nodes, so the block widths split it
    differently): the block width and the row order never change a stored bit or a force sample, nor
    does the pull order (bit 3: loads issued before the node kind is known) with the per-node stores;
    the two store variants (inlet and outlet nodes in a branch of their own, or sharing their wave's
    stores), and the shared-store variant in the two pull orders, are compiled into different branch
    structures, so the compiler's choice of fused multiply-adds can differ between them, and their
    results agree to rounding (the FP32 parity bounds) rather than bit for bit."""
    cfg, sdf = case3(precision, W=150)
    shapes = ((64, 0), (128, 0), (256, 0), (64, 1), (64, 2), (64, 4), (128, 5), (256, 6), (64, 8), (128, 9),
              (64, 12), (256, 14))
    outs = {}
    for block, variant in shapes:
        s = Solver3D.create(dict(cfg))
        s.set_sdf(sdf)
        s.init_field()
        s.set_launch(block, variant)
        s.run(30)
        outs[(block, variant)] = (s.read_populations_raw(), s.read_forces_raw()[1], s.read_populations())
        s.destroy()
    report = {}
    for k, v in outs.items():
        ref = outs[(64, 12)] if k[1] & 12 == 12 else outs[(64, 4)] if k[1] & 4 else outs[(64, 0)]
        report[k] = (int(np.count_nonzero(v[0] != ref[0])), int(np.count_nonzero(v[1] != ref[1])))
    print(report)
    assert all(v == (0, 0) for v in report.values()), report
    tol = TOL32[100] if precision == "fp32" else TOL16[100]
    for other in ((64, 4), (64, 12)):
        a, b = outs[(64, 0)], outs[other]
        assert _err(a[2], b[2]) <= tol
        assert np.abs(a[1] - b[1]).max() <= 1e-4 * np.abs(a[1]).max()


@pytest.mark.parametrize("precision,variant", [("fp32", 0), ("fp32", 4), ("fp16", 1), ("fp32", 12), ("fp16", 9)])
def test_d3q19_zou_he_faces_match_oracle(precision, variant):
    """Hecht-Harting Zou-He inlet and outlet faces (lbm_set_open_faces): GPU against the oracle to the
    bounds of the default planes, in both store variants, with a body two cells from the inlet plane
    (inlet nodes with Bouzidi links go through the boundary list) and one touching the z = 0 face;
    switching back gives the reference's planes bit for bit."""
    cfg, sdf = case3(precision)
    W, H, D = cfg["width"], cfg["height"], cfg["depth"]
    g.add_sphere(sdf, W, H, D, 2.6, 8.0, 6.5, 1.9)
    gpu = Solver3D.create(dict(cfg))
    o32 = OracleSolver3D(dict(cfg), 32)
    o64 = OracleSolver3D(dict(cfg), 64)
    for s in (gpu, o32, o64):
        s.set_sdf(sdf)
        s.init_field()
        s.set_open_faces("zou-he")
    gpu.set_launch(64, variant)
    np.testing.assert_array_equal(gpu.read_flags(), o32.read_flags())
    tol = {"fp32": TOL32, "fp16": TOL16}[precision]
    done = 0
    for n in (1, 10, 100):
        gpu.run(n - done)
        o32.steps(n - done)
        o64.steps(n - done)
        done = n
        e = _err(gpu.read_populations(), o32.read_populations())
        print(f"zou-he {precision} v{variant} step {n}: |gpu-o32| {e:.3e} |gpu-o64| "
              f"{_err(gpu.read_populations(), o64.read_populations()):.3e}")
        assert e <= tol[n], (n, e)
    m = gpu.read_macro().reshape(D, H, W, 4)
    assert np.abs(m[:, :, 0, 3] - 1).max() > 1e-5  # the inlet density floats: not the equilibrium plane
    sg, vg = gpu.read_forces_raw()
    so, vo = o32.read_forces_raw()
    np.testing.assert_array_equal(sg, so)
    assert np.abs(vg - vo).max() / np.abs(vo).max() < {"fp32": 1e-4, "fp16": 2e-3}[precision]
    # back to the reference's planes: the same bits as a solver that never switched
    ref = Solver3D.create(dict(cfg))
    for s in (gpu, ref):
        s.set_sdf(sdf)
        s.init_field()
        s.set_launch(64, variant)
    gpu.set_open_faces("equilibrium")
    for s in (gpu, ref):
        s.run(20)
    np.testing.assert_array_equal(gpu.read_populations_raw(), ref.read_populations_raw())
    for s in (gpu, ref, o32, o64):
        s.destroy()


@pytest.mark.parametrize("precision,smag", [("fp32", 0.0), ("fp32", 0.12), ("fp16", 0.0), ("bf16", 0.0)])
def test_d3q27_fields_match_oracle(precision, smag):
    """D3Q27 (velocity_set 27): GPU against the oracle's D3Q27 on the parity tunnel plus a body on a slip
    edge (corner directions reflected through both faces meet a solid node), flags bit-exact,
    populations within the D3Q19 bounds, force samples at the same steps."""
    cfg, sdf = case3(precision, smag=smag)
    cfg["velocitySet"] = 27
    W, H, D = cfg["width"], cfg["height"], cfg["depth"]
    g.add_sphere(sdf, W, H, D, 12, 0.3, D - 0.8, 1.5)
    gpu = Solver3D.create(dict(cfg))
    o32 = OracleSolver3D(dict(cfg), 32)
    assert gpu.Q == 27 and o32.Q == 27
    for s in (gpu, o32):
        s.set_sdf(sdf)
        s.init_field()
    np.testing.assert_array_equal(gpu.read_flags(), o32.read_flags())
    tol = {"fp32": TOL32, "fp16": TOL16, "bf16": TOLBF}[precision]
    done = 0
    for n in (1, 10, 100):
        gpu.run(n - done)
        o32.steps(n - done)
        done = n
        e = _err(gpu.read_populations(), o32.read_populations())
        print(f"d3q27 {precision} smag {smag} step {n}: |gpu-o32| {e:.3e}")
        assert np.isfinite(gpu.read_populations()).all()
        assert e <= tol[n], (n, e)
    sg, vg = gpu.read_forces_raw()
    so, vo = o32.read_forces_raw()
    np.testing.assert_array_equal(sg, so)
    assert np.abs(vg - vo).max() / np.abs(vo).max() < {"fp32": 1e-4, "fp16": 2e-3, "bf16": 2e-2}[precision]
    m = gpu.read_macro().reshape(D, H, W, 4)
    mo = o32.read_macro().reshape(D, H, W, 4)
    assert np.abs(m - mo).max() < {"fp32": 1e-6, "fp16": 1e-3, "bf16": 5e-3}[precision]
    for s in (gpu, o32):
        s.destroy()


def test_d3q27_uniform_flow_and_unsupported_paths():
    """Without a body the inflow equilibrium is a fixed point of the D3Q27 step on the GPU too; the
    paths written for D3Q19 only (Zou-He faces, the store and pull-order variants, other block widths)
    refuse D3Q27."""
    from cells_amd.lbm import LbmError

    cfg = dict(width=40, height=18, depth=14, precision="fp32", tau=0.53, uIn=0.07, spongeFraction=0.3, spongeTau=1,
               absorb=0.05, velocitySet=27, forceEvery=5)
    s = Solver3D.create(cfg)
    s.init_field()
    f0 = s.read_populations()
    s.run(60)
    assert np.abs(s.read_populations() - f0).max() < 1e-6
    assert len(s.read_forces()) == 12
    with pytest.raises(LbmError):
        s.set_open_faces("zou-he")
    with pytest.raises(LbmError):
        s.set_launch(64, 4)
    with pytest.raises(LbmError):
        s.set_launch(128, 1)
    s.set_launch(64, 2)
    assert s.launch_info() == (64, 2)
    s.destroy()


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_d3q27_row_orders_give_identical_bits(precision):
    """The D3Q27 row orders the autotuner picks from (rows walked bottom-up, or top-down on alternate
    steps): cells are independent within a step and the force partials are indexed by the cell's wave,
    so populations and force samples are the same bits in every order; the autotuner restores the
    state it times on."""
    cfg, sdf = case3(precision, W=150)
    cfg["velocitySet"] = 27
    W, H, D = cfg["width"], cfg["height"], cfg["depth"]
    g.add_sphere(sdf, W, H, D, 12, 0.3, D - 0.8, 1.5)
    outs = {}
    for variant in (0, 1, 2, "auto"):
        s = Solver3D.create(dict(cfg))
        s.set_sdf(sdf)
        s.init_field()
        early = (np.zeros(0, np.int64), np.zeros((0, 3)))
        if variant == "auto":
            s.run(7)
            f7 = s.read_populations_raw()
            early = s.read_forces_raw()
            s.autotune(5)
            np.testing.assert_array_equal(s.read_populations_raw(), f7)
            assert len(s.read_forces_raw()[0]) == 0  # the trial steps leave no samples behind
            assert s.launch_info()[1] in (0, 1, 2)
            s.run(23)
        else:
            s.set_launch(s.launch_info()[0], variant)
            s.run(30)
        late = s.read_forces_raw()
        outs[variant] = (s.read_populations_raw(),
                         (np.concatenate([early[0], late[0]]), np.concatenate([early[1], late[1]])))
        s.destroy()
    for k, v in outs.items():
        np.testing.assert_array_equal(v[0], outs[0][0], err_msg=str(k))
        np.testing.assert_array_equal(v[1][0], outs[0][1][0], err_msg=str(k))
        np.testing.assert_array_equal(v[1][1], outs[0][1][1], err_msg=str(k))


@pytest.mark.parametrize("block", [64, 256])
def test_force_reduction_across_many_blocks(block):
    """The multi-block path of the force reduction (reduce.hpp force_sample: every wave publishes an
    FP64 partial with device-scope stores, waits for them to land, and the last block to bump the
    completion counter sums all partials in wave order): a sphere of radius 14 has a few thousand
    body-link nodes, i.e. 13 to 54 blocks of them at 256 and 64 threads per block. Samples equal the oracle's to the FP32
    bounds of the other force checks at every step, and two solvers give the same bits at every
    sample (a partial read before it landed, or a counter that did not reset, would change a sum or
    drop a sample). The small parity cases fit one block and use the one-block path."""
    W, H, D = 64, 48, 48
    cfg = dict(width=W, height=H, depth=D, precision="fp32", tau=0.56, uIn=0.05, spongeFraction=0.15,
               spongeTau=1, absorb=0.02, smagorinsky=0.0)
    sdf = g.empty_sdf3(W, H, D)
    g.add_sphere(sdf, W, H, D, 24.3, 23.6, 24.2, 14.0)
    sims = [Solver3D.create(dict(cfg)) for _ in range(2)]
    o32 = OracleSolver3D(dict(cfg), 32)
    for s in sims + [o32]:
        s.set_sdf(sdf)
        s.init_field()
    for s in sims:
        s.set_launch(block, 0)
    n_list = sims[0].node_counts()[1]
    assert n_list > 10 * block, n_list  # many blocks of boundary-list nodes
    for s in sims:
        s.run(40)
    o32.steps(40)
    (sa, va), (sb, vb) = sims[0].read_forces_raw(), sims[1].read_forces_raw()
    so, vo = o32.read_forces_raw()
    np.testing.assert_array_equal(sa, so)
    np.testing.assert_array_equal(sa, sb)
    np.testing.assert_array_equal(va.view(np.uint64), vb.view(np.uint64))
    rel = np.abs(va - vo).max() / np.abs(vo).max()
    print(f"block {block}: {n_list} list nodes, {len(sa)} samples, max rel force diff vs oracle {rel:.2e}")
    assert rel < 1e-4
    for s in sims:
        s.destroy()
    o32.destroy()
