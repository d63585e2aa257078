"""Discrete adjoint on the GPU: the gradient of a drag objective with respect to the SDF against the
FP64 torch reference (reverse mode through the same discrete step, tests/adjoint_ref.py), against
central finite differences of the GPU forward (h = 1e-3 in the shape parameters), the
checkpoint-recompute path against the fully recorded one, and the autograd-Function wrapper."""
import math

import numpy as np
import pytest

from cells_amd import geometry as g
from cells_amd import shapes
from cells_amd.adjoint import ShapeAdjoint, TunnelObjective, parameter_gradient
from cells_amd.lbm import X_OPEN, Y_SLIP, Solver, Solver3D

pytestmark = pytest.mark.gpu


def tunnel2(W, H, u, tau=0.56):
    return dict(width=W, height=H, left=X_OPEN, right=X_OPEN, bottom=Y_SLIP, top=Y_SLIP, tau=tau, uIn=u,
                spongeFraction=0.15, spongeTau=1, absorb=0.02, forceEvery=1)


def test_adjoint2d_matches_fp64_reference():
    from tests.adjoint_ref import reference_gradient

    W, H, u, L = 64, 24, 0.05, 8.0
    steps, window, ramp = 200, 80, 60
    cfg = tunnel2(W, H, u)
    sdf = g.empty_sdf(W, H)
    g.add_circle(sdf, W, H, 16, 11.5, 4)
    print(f"adjoint2d vs FP64: {W}x{H}, {steps} steps", flush=True)
    s = Solver.create(dict(cfg))
    adj = ShapeAdjoint(s, steps, window, (1.0, 0.0), 2 / (u * u * L), checkpoint_every=50,
                       ramp=("smoothstep", u, ramp), record_budget_bytes=1 << 40)
    J = adj.forward(sdf)
    gs = adj.backward()
    info = adj.info()
    assert info["stored_segments"] == info["segments"] == 4
    J_ref, g_ref = reference_gradient(cfg, sdf, steps, window, u, ramp, L, segment=50)
    assert abs(J - J_ref) <= 1e-4 * abs(J_ref)
    err = np.linalg.norm(gs - g_ref) / np.linalg.norm(g_ref)
    print(f"J {J:.8f} ref {J_ref:.8f}; |g - g_ref| / |g_ref| = {err:.2e}; nonzero {np.count_nonzero(g_ref)}")
    assert err < 2e-3
    assert np.count_nonzero(gs) == np.count_nonzero(g_ref)
    # the checkpoint-recompute path (no recorded segment) gives the same result bit for bit
    adj2 = ShapeAdjoint(s, steps, window, (1.0, 0.0), 2 / (u * u * L), checkpoint_every=50,
                        ramp=("smoothstep", u, ramp), record_budget_bytes=0)
    assert adj2.forward(sdf) == J
    assert adj2.info()["stored_segments"] == 0
    np.testing.assert_array_equal(adj2.backward(), gs)
    # the seed scales the gradient
    np.testing.assert_allclose(adj2.backward(seed=-2.5), -2.5 * gs, rtol=1e-5, atol=1e-6 * np.abs(gs).max())
    s.destroy()


def test_adjoint2d_naca_bumps_match_finite_differences():
    W, H, u = 128, 48, 0.05
    chord, alpha = 24.0, 5.0
    steps, window, ramp = 600, 200, 200
    L = chord
    cfg = tunnel2(W, H, u, tau=0.58)
    theta = np.array([alpha, 0.0, 0.0, 0.0, 0.0])

    def sdf_of(th):
        return shapes.naca_bumps(W, H, 32.0, 23.5, chord, th[0], th[1:])

    s = Solver.create(dict(cfg))
    adj = ShapeAdjoint(s, steps, window, (1.0, 0.0), 2 / (u * u * L), checkpoint_every=100,
                       ramp=("smoothstep", u, ramp))
    sdf0, dsdf = sdf_of(theta)
    J0 = adj.forward(sdf0)
    grad = parameter_gradient(adj.backward(), dsdf)
    h = 1e-3
    fd = np.zeros_like(theta)
    for p in range(len(theta)):
        e = np.zeros_like(theta)
        e[p] = h
        fd[p] = (adj.forward(sdf_of(theta + e)[0]) - adj.forward(sdf_of(theta - e)[0])) / (2 * h)
    print(f"J {J0:.6f}\nadjoint {grad}\nFD      {fd}")
    assert np.all(np.isfinite(grad))
    for p in range(len(theta)):
        assert abs(grad[p] - fd[p]) <= 0.02 * abs(fd[p]) + 1e-4 * np.abs(fd).max(), p
    s.destroy()


def test_adjoint3d_sphere_harmonics_match_finite_differences():
    """FP32 and FP16-storage forward passes, FP32 adjoint: both gradients against central differences
    of the FP32 forward (the FP16 map itself is too rough at h = 1e-3 for differences: its rounding
    moves J by about 1e-4 between neighbouring shapes)."""
    W, H, D, u = 48, 24, 24, 0.05
    r = 5.0
    steps, window = 300, 100
    area = math.pi * r * r
    coeffs = np.zeros(16)

    def sdf_of(b):
        return shapes.sphere_sh(W, H, D, (12.0, 11.5, 11.5), r, b)

    def make(precision):
        cfg = dict(width=W, height=H, depth=D, precision=precision, tau=0.56, uIn=u, spongeFraction=0.15,
                   spongeTau=1, absorb=0.02, forceEvery=1)
        s = Solver3D.create(cfg)
        return s, ShapeAdjoint(s, steps, window, (1.0, 0.0, 0.0), 2 / (u * u * area), checkpoint_every=50,
                               ramp=("smoothstep", u, 100))

    sdf0, dsdf = sdf_of(coeffs)
    grads = {}
    fd = {}
    for precision in ("fp32", "fp16"):
        s, adj = make(precision)
        J0 = adj.forward(sdf0)
        grads[precision] = parameter_gradient(adj.backward(), dsdf)
        # the forward run does not depend on what ran before it (here: a backward pass)
        assert adj.forward(sdf0) == J0
        if precision == "fp32":
            check = [0, 1, 3, 6, 9]
            fd = {}
            h = 1e-3
            for p in check:
                e = np.zeros(16)
                e[p] = h
                fd[p] = (adj.forward(sdf_of(coeffs + e)[0]) - adj.forward(sdf_of(coeffs - e)[0])) / (2 * h)
        s.destroy()
    scale = max(abs(v) for v in fd.values())
    for precision, tol in (("fp32", 0.01), ("fp16", 0.02)):
        g = grads[precision]
        print(f"{precision}: adjoint {[round(g[p], 5) for p in check]}; FD(fp32) {[round(fd[p], 5) for p in check]}")
        for p in check:
            assert abs(g[p] - fd[p]) <= tol * abs(fd[p]) + 2e-3 * scale, (precision, p)
    # the FP64 reference of the same objective: central differences of the FP64 oracle, all 16
    # coefficients, h = 1e-4, no node flips between fluid and solid (tests/golden/make_adjoint_fd_fp64.py)
    ref = _fp64_fd("sphere_48x24x24")
    assert sum(ref["flips_h0.0001"]) == 0
    f64 = np.array(ref["fd_h0.0001"])
    big = np.abs(f64) >= 0.01 * np.abs(f64).max()
    for precision, tol in (("fp32", 0.01), ("fp16", 0.02)):
        g = grads[precision]
        rel = np.abs(g[big] - f64[big]) / np.abs(f64[big])
        print(f"{precision} adjoint vs FP64 FD: max rel {rel.max():.2e} on {int(big.sum())} components; "
              f"largest error elsewhere {np.abs(g[~big] - f64[~big]).max():.1e}")
        assert rel.max() < tol
        assert np.abs(g[~big] - f64[~big]).max() <= 2e-3 * np.abs(f64).max()


def _fp64_fd(case):
    import json
    import os

    return json.load(open(os.path.join(os.path.dirname(__file__), "golden", "adjoint_fd_fp64.json")))[case]


@pytest.mark.parametrize("nd", [2, 3])
def test_memory_plans_give_identical_gradients(nd):
    """Every memory plan (all segments recorded, some, none; odd and even numbers recomputed, the
    forward pass's recording of the segment below the recorded ones used or not) and a repeated
    backward call give the all-recorded gradient bit for bit."""
    u = 0.05
    if nd == 2:
        W, H, steps, K = 64, 24, 250, 50
        s = Solver.create(tunnel2(W, H, u))
        sdf = g.empty_sdf(W, H)
        g.add_circle(sdf, W, H, 16, 11.5, 4)
        direction = (1.0, 0.0)
    else:
        W, H, D, steps, K = 48, 24, 24, 250, 50
        s = Solver3D.create(dict(width=W, height=H, depth=D, precision="fp16", tau=0.56, uIn=u, spongeFraction=0.15,
                                 spongeTau=1, absorb=0.02, forceEvery=1))
        sdf = shapes.sphere_sh(W, H, D, (12.0, 11.5, 11.5), 5.0, np.zeros(16))[0]
        direction = (1.0, 0.0, 0.0)
    ref = None
    stored = []
    for budget in (1 << 40, 0, 1, 2, 3):
        nbytes = budget if budget in (0, 1 << 40) else budget * (K * s.N * (16 if nd == 2 else 8) + K * 4096) + 1
        adj = ShapeAdjoint(s, steps, 100, direction, 1.0, checkpoint_every=K, ramp=("smoothstep", u, 60),
                           record_budget_bytes=nbytes)
        J = adj.forward(sdf)
        grads = [adj.backward(), adj.backward()]
        stored.append(adj.info()["stored_segments"])
        if ref is None:
            ref = (J, grads[0])
        print(f"d{nd} budget {nbytes}: {stored[-1]} of {adj.info()['segments']} segments recorded", flush=True)
        assert J == ref[0]
        for gr in grads:
            np.testing.assert_array_equal(gr, ref[1])
        adj.release()
    assert stored[0] == 5 and stored[1] == 0 and len(set(stored)) >= 4, stored
    s.destroy()


def test_autograd_function_semantics():
    import torch

    W, H, u = 64, 24, 0.05
    cfg = tunnel2(W, H, u)
    sdf = g.empty_sdf(W, H)
    g.add_circle(sdf, W, H, 16, 11.5, 4)
    s = Solver.create(dict(cfg))
    adj = ShapeAdjoint(s, 120, 40, (1.0, 0.0), 1.0, checkpoint_every=40, ramp=("smoothstep", u, 40))
    t = torch.tensor(sdf, dtype=torch.float32, requires_grad=True)
    J = TunnelObjective.apply(t, adj)
    (3.0 * J).backward()
    direct = adj.backward(seed=3.0)
    np.testing.assert_array_equal(t.grad.numpy(), direct.astype(np.float32))
    assert float(J) == adj.forward(sdf)
    s.destroy()


def test_adjoint2d_naca_384x128_matches_fp64_fixture():
    """The 2D shape-gradient case at full size (384 x 128, 6000 steps, 9 parameters) against the FP64
    reverse-mode reference computed on the CPU by tests/golden/make_adjoint_naca_fp64.py."""
    import json
    import os

    import bench

    ref = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "adjoint_naca_384x128_fp64.json")))
    out = bench.adjoint2d_case(fd=False)
    g, g_ref = np.array(out["grad"]), np.array(ref["grad_theta"])
    err = np.abs(g - g_ref) / np.abs(g_ref)
    print(f"J {out['J']:.9f} ref {ref['J']:.9f}; max relative gradient error {err.max():.2e}")
    assert abs(out["J"] - ref["J"]) <= 1e-4 * abs(ref["J"])
    assert err.max() < 1e-3
    # and against central differences of the FP64 oracle (h = 1e-4; no node flips at +-h; at h = 1e-3
    # the difference quotient of a_1 is 2% off its own h -> 0 limit)
    fd = _fp64_fd("naca_384x128")
    assert sum(fd["flips_h0.0001"]) == 0 and sum(fd["flips_h0.001"]) == 0
    rel = np.abs(g - np.array(fd["fd_h0.0001"])) / np.abs(np.array(fd["fd_h0.0001"]))
    print(f"vs FP64 central differences (h 1e-4): max rel {rel.max():.2e}; J {ref['J']:.12f} (reverse mode) "
          f"{fd['J']:.12f} (oracle)")
    assert abs(fd["J"] - ref["J"]) <= 1e-9 * abs(ref["J"])  # two FP64 references of one objective
    assert rel.max() < 0.01
    # at the configured h = 1e-3: within 1% on every parameter whose difference quotient does not
    # straddle a Bouzidi branch switch (only a_1 does: one link's q crosses 1/2 between theta +- h e_1,
    # tests/test_adjoint_fd_diagnostics.py)
    from cells_amd import shapes
    from cells_amd.adjoint import bouzidi_branch_switches

    theta = np.array([5.0] + [0.0] * 8)
    switches = []
    for p in range(9):
        e = np.zeros(9)
        e[p] = 1e-3
        sp, sm = (shapes.naca_bumps(384, 128, 96.0, 63.5, 64.0, t[0], t[1:])[0] for t in (theta + e, theta - e))
        switches.append(bouzidi_branch_switches(sp, sm, 384, 128))
    assert switches == [0, 1, 0, 0, 0, 0, 0, 0, 0]
    smooth = [p for p in range(9) if switches[p] == 0]
    f3 = np.array(fd["fd_h0.001"])
    rel3 = np.abs(g[smooth] - f3[smooth]) / np.abs(f3[smooth])
    print(f"vs FP64 central differences (h 1e-3) on the {len(smooth)} parameters without a branch switch: "
          f"max rel {rel3.max():.2e}")
    assert rel3.max() < 0.01
    # the configured memory plan (recompute from checkpoints) and the recorded one agree bit for bit
    assert out["info"]["stored_segments"] == 0 and out["recorded"]["info"]["stored_segments"] == 30
    print(f"adjoint / forward: {out['ratio']:.2f} (recompute), {out['recorded']['ratio']:.2f} (recorded)")


def test_adjoint3d_sphere_256x128x128_fp16():
    """The 3D shape-gradient case at full size (256 x 128 x 128, FP16-storage forward, FP32 adjoint,
    3000 steps, checkpoints every 100 steps): recomputed and recorded segments give the same gradient
    bit for bit, and the gradient of four coefficients that the symmetry leaves non-zero (l = 0, 1, 2) matches
    central differences of the FP32 forward (h = 1e-3) within 1%."""
    import bench

    out = bench.adjoint3d_case(fd_coeffs=(0, 3, 6, 8))
    print({k: out[k] for k in ("ratio", "max_rel_err_vs_fd", "max_rel_err_fp32_adjoint_vs_fd")},
          out["recorded"]["ratio"])
    assert out["recorded"]["same_gradient_bits"]
    assert out["info"]["stored_segments"] == 0 and out["info"]["segments"] == 30
    assert out["max_rel_err_vs_fd"] < 0.01 and out["max_rel_err_fp32_adjoint_vs_fd"] < 0.01



def test_adjoint3d_open_planes_against_fp64_differences():
    """A short tunnel with no sponge and the body two cells from the outlet plane, so that the
    reverse step's rules for the x planes next to the open ends (inlet plane mu = 0, no reader of the
    c_x < 0 populations at x = 1, the outlet's copy rule reading x = W - 2, the outlet equilibrium's
    transpose at x = W - 1) all reach the body within the run: every spherical-harmonic coefficient's
    gradient against central differences of the FP64 oracle (h = 1e-4, computed here), and the open
    planes are streamed, not listed."""
    from oracle.oracle import OracleSolver3D

    W, H, D, u, r = 14, 10, 10, 0.05, 2.5
    steps, window = 60, 30
    c = (8.5, 4.7, 4.6)
    cfg = dict(width=W, height=H, depth=D, precision="fp32", tau=0.6, uIn=u, spongeFraction=0.0, absorb=0.0,
               forceEvery=1)
    scale = 2 / (u * u * math.pi * r * r)
    coeffs = np.zeros(16)

    def sdf_of(b):
        return shapes.sphere_sh(W, H, D, c, r, b)

    def oracle_J(sdf):
        o = OracleSolver3D(dict(cfg), 64)
        o.set_sdf(sdf)
        o.set_inlet_ramp("smoothstep", u, 20)
        o.set_inlet(0.0)
        o.init_field()
        o.steps(steps)
        st, v = o.read_forces_raw()
        o.destroy()
        return scale * v[st > steps - window, 0].mean()

    s = Solver3D.create(dict(cfg))
    adj = ShapeAdjoint(s, steps, window, (1.0, 0.0, 0.0), scale, checkpoint_every=20, ramp=("smoothstep", u, 20))
    sdf0, dsdf = sdf_of(coeffs)
    J = adj.forward(sdf0)
    grad = parameter_gradient(adj.backward(), dsdf)
    info = adj.info()
    J64 = oracle_J(sdf0)
    h = 1e-4
    fd = np.array([(oracle_J(sdf_of(coeffs + h * e)[0]) - oracle_J(sdf_of(coeffs - h * e)[0])) / (2 * h)
                   for e in np.eye(16)])
    big = np.abs(fd) >= 0.01 * np.abs(fd).max()
    rel = np.abs(grad[big] - fd[big]) / np.abs(fd[big])
    print(f"J {J:.7f} (FP64 {J64:.7f}); {info}; adjoint vs FP64 differences: max rel {rel.max():.2e} on "
          f"{int(big.sum())} components, largest error elsewhere {np.abs(grad - fd)[~big].max():.1e}")
    assert abs(J - J64) <= 1e-5 * abs(J64)
    assert rel.max() < 2e-3
    assert np.abs(grad - fd)[~big].max() <= 1e-3 * np.abs(fd).max()
    # the open planes and their neighbours (4 H D nodes) are streamed; the list holds the body's nodes
    assert info["list_nodes"] < 4 * H * D
    s.destroy()
