"""The population stride (lbm_set_population_stride, and the stride stage of lbm_autotune; D3Q19 and
D2Q9): the device buffers may hold each population at a padded stride, which changes which memory
channels a node's streams share and nothing a caller can see. Every test here runs a padded solver
beside an unpadded one on the same inputs and requires the same bits: populations, force samples,
flags, macro fields, refilled cells after a geometry change, batched tunnels, and the adjoint's
gradient."""
from __future__ import annotations

import numpy as np
import pytest

from cells_amd import geometry as g
from cells_amd import shapes, sweep
from cells_amd.adjoint import ShapeAdjoint
from cells_amd.lbm import Batch, LbmError, Solver, Solver3D
from tests.parity_cases import CASES2, case3

pytestmark = pytest.mark.gpu


def _padded(N, pad):
    return (N + 1023) // 1024 * 1024 + pad


def _make(cfg, sdf, launch=None, faces=None):
    s = Solver3D.create(dict(cfg))
    s.set_sdf(sdf)
    if faces:
        s.set_open_faces(faces)
    s.init_field()
    if launch:
        s.set_launch(*launch)
    return s


def _state(s):
    s.encode_macro()
    return dict(pop=s.read_populations_raw(), forces=s.read_forces_raw(), flags=s.read_flags(), macro=s.read_macro())


def _same(a, b):
    assert np.array_equal(a["pop"], b["pop"]), int(np.count_nonzero(a["pop"] != b["pop"]))
    assert np.array_equal(a["forces"][0], b["forces"][0])
    assert np.array_equal(a["forces"][1], b["forces"][1])
    assert np.array_equal(a["flags"], b["flags"])
    assert np.array_equal(a["macro"], b["macro"])


@pytest.mark.parametrize("precision", ["fp32", "fp16", "bf16", "fp16sr"])
def test_padded_stride_gives_identical_bits(precision):
    """A row of 150 nodes (blocks split it unevenly), N = 144,000 (not a multiple of 64), two bodies
    (Bouzidi links, a slip face through a solid node): the stride moved three times during the run."""
    cfg, sdf = case3(precision, W=150, H=40, D=24)
    ref = _make(cfg, sdf, launch=(64, 1))
    pad = _make(cfg, sdf, launch=(64, 1))
    N = pad.N
    assert pad.population_stride() == N
    s1, s2 = _padded(N, 4096), _padded(N, 17 * 4096)
    pad.set_population_stride(s1)
    assert pad.population_stride() == s1
    ref.run(25)
    pad.run(25)
    pad.set_population_stride(s2)
    ref.run(20)
    pad.run(20)
    pad.set_population_stride(s1)
    ref.run(15)
    pad.run(15)
    assert pad.population_stride() == s1
    assert pad.node_counts() == ref.node_counts()
    _same(_state(ref), _state(pad))
    pad.set_population_stride(0)
    assert pad.population_stride() == N
    ref.run(5)
    pad.run(5)
    _same(_state(ref), _state(pad))
    ref.destroy()
    pad.destroy()


@pytest.mark.parametrize("precision,smag", [("fp32", 0.0), ("fp16", 0.0), ("bf16", 0.12)])
def test_d3q27_padded_stride_gives_identical_bits(precision, smag):
    """D3Q27 resolves its boundary pulls in the kernel (mirrored rows by arithmetic, nodes next to a
    body per direction): both read the stride."""
    cfg, sdf = case3(precision, W=150, H=40, D=24, smag=smag)
    cfg = dict(cfg, velocitySet=27)
    ref = _make(cfg, sdf, launch=(192, 1))
    pad = _make(cfg, sdf, launch=(192, 1))
    N = pad.N
    pad.set_population_stride(_padded(N, 4096))
    ref.run(20)
    pad.run(20)
    pad.set_population_stride(_padded(N, 17 * 4096))
    ref.run(20)
    pad.run(20)
    _same(_state(ref), _state(pad))
    ref.destroy()
    pad.destroy()


@pytest.mark.parametrize("faces,smag,variant", [("equilibrium", 0.0, 0), ("equilibrium", 0.0, 14),
                                                ("zou-he", 0.0, 5), ("zou-he", 0.0, 12), ("equilibrium", 0.12, 2)])
def test_padded_stride_kernel_variants(faces, smag, variant):
    """Each kernel instantiation reads the stride: per-node and shared stores, both pull orders, the
    Zou-He faces and the subgrid model."""
    cfg, sdf = case3("fp32", W=150, H=40, D=24, smag=smag)
    ref = _make(cfg, sdf, launch=(128, variant), faces=faces)
    pad = _make(cfg, sdf, launch=(128, variant), faces=faces)
    pad.set_population_stride(_padded(pad.N, 17 * 4096))
    ref.run(30)
    pad.run(30)
    _same(_state(ref), _state(pad))
    ref.destroy()
    pad.destroy()


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_padded_stride_geometry_change_refills(precision):
    """setSdf on a developed flow at a padded stride: the flag kernel refills uncovered cells into the
    padded buffers and the boundary tables are rebuilt with the stride."""
    cfg, sdf = case3(precision, W=150, H=40, D=24)
    W, H, D = 150, 40, 24
    ref = _make(cfg, sdf, launch=(256, 2))
    pad = _make(cfg, sdf, launch=(256, 2))
    pad.set_population_stride(_padded(pad.N, 4096))
    ref.run(30)
    pad.run(30)
    sdf2 = g.empty_sdf3(W, H, D)
    g.add_sphere(sdf2, W, H, D, 11.0, 7.5, 6.5, 3.5)
    g.add_sphere(sdf2, W, H, D, 40, 30, 0.5, 4.2)
    for s in (ref, pad):
        s.set_sdf(sdf2)
    _same(_state(ref), _state(pad))
    ref.run(15)
    pad.run(15)
    _same(_state(ref), _state(pad))
    ref.destroy()
    pad.destroy()


def test_autotune_times_strides_and_restores_state():
    """N = 144,000 >= 65,536: the autotuner's stride stage runs. Its pick is one of the candidates,
    the state it leaves is the state it found (at whatever stride it chose), a second solver of the
    same configuration gets the same pick from the process cache, and the run continues bit for bit
    like an unpadded solver pinned to the same launch shape."""
    cfg, sdf = case3("fp16", W=150, H=40, D=24)
    a = _make(cfg, sdf)
    a.run(20)
    before = _state(a)
    a.autotune(10)
    N = a.N
    stride = a.population_stride()
    assert stride in (N, _padded(N, 4096), _padded(N, 17 * 4096))
    after = _state(a)
    assert np.array_equal(before["pop"], after["pop"])
    assert np.array_equal(before["macro"], after["macro"])
    block, variant = a.launch_info()
    # b repeats a's launch history at stride N: the default shape for the first 20 steps, then the
    # pick (the store variants agree with each other only to rounding)
    b = _make(cfg, sdf)
    b.run(20)
    b.set_launch(block, variant)
    b.read_forces_raw()
    a.run(25)
    b.run(25)
    _same(_state(a), _state(b))
    c = _make(cfg, sdf)
    c.autotune(10)
    assert c.population_stride() == stride and c.launch_info() == (block, variant)
    for s in (a, b, c):
        s.destroy()


def test_adjoint_gradient_identical_at_padded_stride():
    """The adjoint's checkpoints hold whole buffers at the solver's stride and its tables decode the
    forward gather indices through it: J and dJ/dsdf are the same bits with a padded forward state.
    A stride change is refused while the adjoint state exists."""
    W, H, D, u, r = 48, 24, 24, 0.05, 5.0
    sdf0, _ = shapes.sphere_sh(W, H, D, (12.0, 11.5, 11.5), r, np.zeros(16))
    out = {}
    for padded in (False, True):
        cfg = dict(width=W, height=H, depth=D, precision="fp16", tau=0.56, uIn=u, spongeFraction=0.15,
                   spongeTau=1, absorb=0.02, forceEvery=1)
        s = Solver3D.create(cfg)
        if padded:
            s.set_population_stride(_padded(s.N, 17 * 4096))
        adj = ShapeAdjoint(s, 150, 50, (1.0, 0.0, 0.0), 1.0, checkpoint_every=50, ramp=("smoothstep", u, 50))
        J = adj.forward(sdf0)
        out[padded] = (J, adj.backward())
        if padded:
            with pytest.raises(LbmError):
                s.set_population_stride(0)
            adj.release()
            s.set_population_stride(0)
            assert s.population_stride() == s.N
        s.destroy()
    assert out[False][0] == out[True][0]
    assert np.array_equal(out[False][1], out[True][1])


def test_stride_argument_checks():
    cfg, sdf = case3("fp32", W=150, H=40, D=24)
    s = _make(cfg, sdf)
    N = s.N
    for bad in (N - 64, _padded(N, 4096) + 32, 1 << 30):
        with pytest.raises(LbmError):
            s.set_population_stride(bad)
    assert s.population_stride() == N
    s.destroy()
    cfg27 = dict(cfg, velocitySet=27)
    q = _make(cfg27, sdf)
    with pytest.raises(LbmError):
        q.set_population_stride(q.N + 32)
    q.set_population_stride(0)
    assert q.population_stride() == q.N
    q.destroy()
    d2 = Solver.create(dict(width=128, height=64))
    d2.set_population_stride(_padded(d2.N, 4096))
    assert d2.population_stride() == _padded(d2.N, 4096)
    with pytest.raises(LbmError):
        d2.set_population_stride(d2.N - 64)
    d2.destroy()


@pytest.mark.parametrize("name", ["tunnel", "channel", "periodic", "cavity"])
def test_d2q9_padded_stride_gives_identical_bits(name):
    """D2Q9: Zou-He planes and a Bouzidi cylinder, no-slip walls with the subgrid model, periodic seams
    with a body across them and a body force, a moving lid; the stride moved twice during the run."""
    cfg, sdf_fn = CASES2[name]
    W, H = cfg["width"], cfg["height"]

    def make():
        s = Solver.create(dict(cfg))
        if sdf_fn:
            s.set_sdf(sdf_fn(W, H))
        s.init_field()
        s.set_launch(64, 1)
        return s

    ref, pad = make(), make()
    N = pad.N
    pad.set_population_stride(_padded(N, 4096))
    ref.run(20)
    pad.run(20)
    pad.set_population_stride(_padded(N, 17 * 4096))
    ref.run(25)
    pad.run(25)
    assert pad.node_counts() == ref.node_counts()
    _same(_state(ref), _state(pad))
    ref.destroy()
    pad.destroy()


def test_batch_with_padded_tunnels_gives_identical_bits():
    """Batched tunnels whose strides change through the handles the caller kept (without fetching them
    again from the batch): the batch notices the new buffers and tables before its next launch, its
    autotuner saves and restores the padded buffers, and every tunnel stores the bits of the unpadded
    batch."""
    cases = sweep.sweep_cases(n=5, seed=7, sizes=(64, 96, 160))

    def run(pads):
        b = Batch.create([c["config"] for c in cases])
        for i, c in enumerate(cases):
            b.tunnel(i).set_sdf(sweep.case_sdf(c))
            b.tunnel(i).init_field()
        b.reset_clock()
        b.set_launch(256, 2)
        b.run(40)
        for i, t in enumerate(b.tunnels):
            if pads:
                t.set_population_stride(_padded(t.N, pads[i % len(pads)]))
        b.autotune(5)
        b.run(70)
        out = [(t.read_populations_raw(), t.read_forces_raw()[1], t.population_stride()) for t in b.tunnels]
        b.destroy()
        return out

    ref = run(())
    got = run((4096, 17 * 4096, 0))
    assert any(g[2] != r[2] for g, r in zip(got, ref))
    for (pg, fg, _), (pr, fr, _) in zip(got, ref):
        np.testing.assert_array_equal(pg, pr)
        np.testing.assert_array_equal(fg, fr)


# ---------------------------------------------------------------- tiles of 64 cells (D3Q19, W % 64 == 0)

@pytest.mark.parametrize("precision", ["fp32", "fp16", "bf16", "fp16sr"])
def test_tiled_layout_gives_identical_bits(precision):
    """Rows of 192 nodes (three tiles each), two bodies, a slip face through a solid node: the layout
    moves from tiles to a padded stride and back during the run."""
    cfg, sdf = case3(precision, W=192, H=40, D=24)
    ref = _make(cfg, sdf, launch=(64, 1))
    til = _make(cfg, sdf, launch=(64, 1))
    til.set_population_layout(True)
    assert til.population_layout() == (1, 64)
    ref.run(25)
    til.run(25)
    _same(_state(ref), _state(til))
    til.set_population_stride(_padded(til.N, 4096))
    assert til.population_layout() == (0, _padded(til.N, 4096))
    ref.run(10)
    til.run(10)
    til.set_population_layout(True)
    ref.run(15)
    til.run(15)
    assert til.node_counts() == ref.node_counts()
    _same(_state(ref), _state(til))
    ref.destroy()
    til.destroy()


@pytest.mark.parametrize("faces,smag,block,variant", [("equilibrium", 0.0, 192, 0), ("equilibrium", 0.0, 64, 14),
                                                      ("zou-he", 0.0, 128, 5), ("zou-he", 0.0, 192, 12),
                                                      ("equilibrium", 0.12, 64, 2), ("equilibrium", 0.0, 128, 9)])
def test_tiled_layout_kernel_variants(faces, smag, block, variant):
    """Every kernel instantiation and block width (64, 128 and 192 threads, so blocks start inside a
    row) on the tiled layout."""
    cfg, sdf = case3("fp32", W=192, H=40, D=24, smag=smag)
    ref = _make(cfg, sdf, launch=(block, variant), faces=faces)
    til = _make(cfg, sdf, launch=(block, variant), faces=faces)
    til.set_population_layout(True)
    ref.run(30)
    til.run(30)
    _same(_state(ref), _state(til))
    ref.destroy()
    til.destroy()


def test_tiled_layout_geometry_change_refills():
    cfg, sdf = case3("fp16", W=192, H=40, D=24)
    W, H, D = 192, 40, 24
    ref = _make(cfg, sdf, launch=(192, 2))
    til = _make(cfg, sdf, launch=(192, 2))
    til.set_population_layout(True)
    ref.run(30)
    til.run(30)
    sdf2 = g.empty_sdf3(W, H, D)
    g.add_sphere(sdf2, W, H, D, 11.0, 7.5, 6.5, 3.5)
    g.add_sphere(sdf2, W, H, D, 70, 30, 0.5, 4.2)
    for s in (ref, til):
        s.set_sdf(sdf2)
    _same(_state(ref), _state(til))
    ref.run(15)
    til.run(15)
    _same(_state(ref), _state(til))
    ref.destroy()
    til.destroy()


def test_tiled_layout_adjoint_gradient_identical():
    W, H, D, u, r = 64, 24, 24, 0.05, 5.0
    sdf0, _ = shapes.sphere_sh(W, H, D, (14.0, 11.5, 11.5), r, np.zeros(16))
    out = {}
    for tiled in (False, True):
        cfg = dict(width=W, height=H, depth=D, precision="fp16", tau=0.56, uIn=u, spongeFraction=0.15,
                   spongeTau=1, absorb=0.02, forceEvery=1)
        s = Solver3D.create(cfg)
        if tiled:
            s.set_population_layout(True)
        adj = ShapeAdjoint(s, 150, 50, (1.0, 0.0, 0.0), 1.0, checkpoint_every=50, ramp=("smoothstep", u, 50))
        out[tiled] = (adj.forward(sdf0), adj.backward())
        if tiled:
            with pytest.raises(LbmError):
                s.set_population_layout(False)
            adj.release()
        s.destroy()
    assert out[False][0] == out[True][0]
    assert np.array_equal(out[False][1], out[True][1])


def test_autotune_layouts_include_tiles_and_restore_state():
    cfg, sdf = case3("fp32", W=192, H=40, D=24)
    a = _make(cfg, sdf)
    a.run(20)
    before = _state(a)
    a.autotune(10)
    N = a.N
    assert a.population_layout() in ((0, N), (0, _padded(N, 4096)), (0, _padded(N, 17 * 4096)), (1, 64))
    after = _state(a)
    assert np.array_equal(before["pop"], after["pop"])
    block, variant = a.launch_info()
    b = _make(cfg, sdf)
    b.run(20)
    b.set_launch(block, variant)
    b.read_forces_raw()
    a.run(25)
    b.run(25)
    _same(_state(a), _state(b))
    # from the tiled layout, every candidate again (the guard restores into whichever is current)
    c = _make(cfg, sdf)
    c.set_population_layout(True)
    c.run(20)
    before = _state(c)
    c.autotune(10)
    assert np.array_equal(before["pop"], _state(c)["pop"])
    for s in (a, b, c):
        s.destroy()


@pytest.mark.parametrize("precision,smag", [("fp32", 0.0), ("fp16", 0.0), ("bf16", 0.12)])
def test_d3q27_tiled_layout_gives_identical_bits(precision, smag):
    """D3Q27 in tiles of 64 cells: the fast path's row-uniform bases with per-lane x offsets, the
    per-direction pulls of the nodes next to a body, the stores, refill and macro kernels."""
    cfg, sdf = case3(precision, W=192, H=40, D=24, smag=smag)
    cfg = dict(cfg, velocitySet=27)
    ref = _make(cfg, sdf, launch=(192, 1))
    til = _make(cfg, sdf, launch=(192, 1))
    til.set_population_layout(True)
    ref.run(25)
    til.run(25)
    _same(_state(ref), _state(til))
    sdf2 = g.empty_sdf3(192, 40, 24)
    g.add_sphere(sdf2, 192, 40, 24, 11.0, 7.5, 6.5, 3.5)
    for s in (ref, til):
        s.set_sdf(sdf2)
    ref.run(15)
    til.run(15)
    _same(_state(ref), _state(til))
    ref.destroy()
    til.destroy()


def test_tiled_layout_argument_checks():
    cfg, sdf = case3("fp32", W=150, H=40, D=24)
    s = _make(cfg, sdf)
    with pytest.raises(LbmError):
        s.set_population_layout(True)  # rows of 150 nodes
    s.destroy()
    cfg, sdf = case3("fp32", W=192, H=40, D=24)
    s = _make(cfg, sdf)
    with pytest.raises(LbmError):
        s.set_population_layout(True, 32)
    s.set_population_layout(True, 64)
    assert s.population_layout() == (1, 64)
    s.set_population_layout(False)
    assert s.population_layout() == (0, s.N)
    s.destroy()
    q = _make(dict(cfg, velocitySet=27), sdf)
    q.set_population_layout(True)
    assert q.population_layout() == (1, 64)
    q.destroy()
    d2 = Solver.create(dict(width=100, height=64))
    with pytest.raises(LbmError):
        d2.set_population_layout(True)  # rows of 100 nodes
    d2.destroy()
    d2 = Solver.create(dict(width=128, height=64))
    d2.set_population_layout(True)
    assert d2.population_layout() == (1, 64)
    d2.destroy()


@pytest.mark.parametrize("name", ["tunnel", "channel", "periodic", "cavity"])
def test_d2q9_tiled_layout_gives_identical_bits(name):
    """The 2D parity cases widened to 128 columns (two tiles per row): Zou-He planes and a body, no-slip
    walls with the subgrid model, periodic seams, a moving lid."""
    cfg, sdf_fn = CASES2[name]
    cfg = dict(cfg, width=128)
    W, H = 128, cfg["height"]

    def make():
        s = Solver.create(dict(cfg))
        if sdf_fn:
            s.set_sdf(sdf_fn(W, H))
        s.init_field()
        s.set_launch(64, 1)
        return s

    ref, til = make(), make()
    til.set_population_layout(True)
    ref.run(25)
    til.run(25)
    _same(_state(ref), _state(til))
    til.set_population_stride(_padded(til.N, 4096))
    ref.run(10)
    til.run(10)
    til.set_population_layout(True)
    ref.run(10)
    til.run(10)
    _same(_state(ref), _state(til))
    ref.destroy()
    til.destroy()


def test_batch_tiled_layout_and_its_autotuning():
    """Batched tunnels (rows of 64, 128 and 192 nodes) tiled through the handles the caller kept, and the
    batch autotuner on tiled tunnels (it saves and restores each tunnel in its layout): the same bits
    as the batch with one array per population."""
    cases = sweep.sweep_cases(n=5, seed=11, sizes=(64, 128, 192))

    def run(mode):
        b = Batch.create([c["config"] for c in cases])
        for i, c in enumerate(cases):
            b.tunnel(i).set_sdf(sweep.case_sdf(c))
            b.tunnel(i).init_field()
        b.reset_clock()
        b.set_launch(256, 2)
        b.run(40)
        if mode in ("tiled", "autotune"):
            for t in b.tunnels:
                t.set_population_layout(True)
        if mode == "autotune":
            before = [t.read_populations_raw() for t in b.tunnels]
            b.autotune(5)
            for t, p0 in zip(b.tunnels, before):
                np.testing.assert_array_equal(t.read_populations_raw(), p0)
        b.run(70)
        out = [(t.read_populations_raw(), t.read_forces_raw()[1], t.population_layout()) for t in b.tunnels]
        b.destroy()
        return out

    ref = run("plain")
    for mode in ("tiled", "autotune"):
        got = run(mode)
        assert all(g[2] == (1, 64) for g in got)
        for (pg, fg, _), (pr, fr, _) in zip(got, ref):
            np.testing.assert_array_equal(pg, pr)
            np.testing.assert_array_equal(fg, fr)


def test_d2q9_adjoint_gradient_identical_in_tiles():
    """The 2D adjoint's tables take the forward sources through the layout: J and dJ/dsdf are the same
    bits with the forward state in tiles of 64 cells."""
    from cells_amd.lbm import X_OPEN, Y_SLIP

    W, H, u = 128, 32, 0.05
    sdf = g.empty_sdf(W, H)
    g.add_circle(sdf, W, H, 30, 15.5, 4)
    out = {}
    for tiled in (False, True):
        s = Solver.create(dict(width=W, height=H, left=X_OPEN, right=X_OPEN, bottom=Y_SLIP, top=Y_SLIP, tau=0.56,
                               uIn=u, spongeFraction=0.15, spongeTau=1, absorb=0.02, forceEvery=1))
        if tiled:
            s.set_population_layout(True)
        adj = ShapeAdjoint(s, 200, 80, (1.0, 0.0), 1.0, checkpoint_every=50, ramp=("smoothstep", u, 60))
        out[tiled] = (adj.forward(sdf), adj.backward())
        adj.release()
        s.destroy()
    assert out[False][0] == out[True][0]
    assert np.array_equal(out[False][1], out[True][1])
