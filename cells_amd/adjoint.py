"""Shape gradients through the discrete adjoint of the tunnel step (no counterpart in the reference).

`ShapeAdjoint` wraps the C ABI pair lbm_adjoint_forward / lbm_adjoint_backward of one solver:
forward runs the tunnel from its initial field and returns J = scale * mean over a window of force
samples of direction . F (C_D with direction (1, 0, 0) and scale 2 / (u^2 L)), backward returns
dJ/dsdf at every node. `TunnelObjective` is the same pair as a torch.autograd.Function over the SDF
tensor, and `parameter_gradient` chains dJ/dsdf with the derivatives of a parametrized SDF
(cells_amd/shapes.py).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lbm as _lbm

RAMP_KINDS = {None: 0, "none": 0, "linear": 1, "smoothstep": 2}


class ShapeAdjoint:
    def __init__(self, solver, steps: int, window: int, direction=(1.0, 0.0, 0.0), scale: float = 1.0,
                 checkpoint_every: int = 200, ramp=None, record_budget_bytes: int = 0):
        """ramp = (kind, u_target, ramp_steps) as Solver.set_inlet_ramp, or None for a constant u_in."""
        self.solver = solver
        kind, u, rs = (ramp if ramp is not None else (None, 0.0, 0.0))
        cfg = _lbm.AdjointConfig()
        cfg.steps = int(steps)
        cfg.window_start = int(steps - window)
        d = list(direction) + [0.0] * (3 - len(direction))
        cfg.direction = (C.c_double * 3)(*[float(v) for v in d[:3]])
        cfg.scale = float(scale)
        cfg.checkpoint_every = int(checkpoint_every)
        cfg.ramp_kind = RAMP_KINDS[kind]
        cfg.ramp_u = float(u)
        cfg.ramp_steps = float(rs)
        cfg.record_budget_bytes = int(record_budget_bytes)
        self.cfg = cfg

    def forward(self, sdf=None) -> float:
        if sdf is not None:
            self.solver.set_sdf(np.ascontiguousarray(sdf, dtype=np.float32))
        J = C.c_double()
        _lbm._check(_lbm.load_library().lbm_adjoint_forward(self.solver._h, C.byref(self.cfg), C.byref(J)))
        return float(J.value)

    def backward(self, seed: float = 1.0) -> np.ndarray:
        out = np.zeros(self.solver.N, dtype=np.float64)
        _lbm._check(_lbm.load_library().lbm_adjoint_backward(self.solver._h, float(seed), out.ctypes.data, out.size))
        return out

    def info(self) -> dict:
        v = [C.c_int64() for _ in range(5)]
        _lbm._check(_lbm.load_library().lbm_adjoint_info(self.solver._h, *[C.byref(x) for x in v]))
        return dict(regular_nodes=v[0].value, list_nodes=v[1].value, entries=v[2].value,
                    stored_segments=v[3].value, segments=v[4].value)

    def release(self):
        _lbm._check(_lbm.load_library().lbm_adjoint_release(self.solver._h))


def parameter_gradient(dJ_dsdf: np.ndarray, dsdf_dtheta: np.ndarray) -> np.ndarray:
    """dJ/dtheta = sum over nodes of dJ/dsdf * dsdf/dtheta (dsdf_dtheta is [P, N])."""
    return np.asarray(dsdf_dtheta, dtype=np.float64) @ np.asarray(dJ_dsdf, dtype=np.float64)


_C2 = np.array([[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1], [1, 1], [-1, 1], [-1, -1], [1, -1]])
_C3 = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [1, 1, 0],
                [-1, -1, 0], [1, 0, 1], [-1, 0, -1], [0, 1, 1], [0, -1, -1], [1, -1, 0], [-1, 1, 0], [1, 0, -1],
                [-1, 0, 1], [0, 1, -1], [0, -1, 1]])


def bouzidi_links(sdf, W: int, H: int, D: int = 1):
    """The Bouzidi links of a geometry: (fluid node, direction) pairs whose upstream node x - c_i lies
    on the grid and is solid (solid: sdf < 0 except on the open x = 0 and x = W - 1 planes, as the
    flag kernels decide), with q = clamp(dF / max(dF - dS, 1e-6), 1e-3, 1) in float32, dF and dS the
    SDF at the node and at its upstream node (src/shaders/step.ts:138-145, step3d.ts:149-156).
    Returns {direction: (node indices, q)} for D2Q9 (D = 1) or D3Q19."""
    s = np.asarray(sdf, dtype=np.float32).reshape(D, H, W)
    solid = s < 0
    solid[:, :, 0] = solid[:, :, W - 1] = False
    cs = _C2 if D == 1 else _C3
    out = {}
    for i in range(1, len(cs)):
        cx, cy = int(cs[i][0]), int(cs[i][1])
        cz = int(cs[i][2]) if D > 1 else 0
        # node (x, y, z) with upstream (x - cx, y - cy, z - cz) on the grid
        xs = slice(max(cx, 0), W + min(cx, 0))
        ys = slice(max(cy, 0), H + min(cy, 0))
        zs = slice(max(cz, 0), D + min(cz, 0))
        us = (slice(zs.start - cz, zs.stop - cz), slice(ys.start - cy, ys.stop - cy), slice(xs.start - cx, xs.stop - cx))
        node_fluid = ~solid[zs, ys, xs]
        up_solid = solid[us]
        link = node_fluid & up_solid
        dF, dS = s[zs, ys, xs][link], s[us][link]
        q = np.clip(dF / np.maximum(dF - dS, np.float32(1e-6)), np.float32(1e-3), np.float32(1.0))
        z, y, x = np.nonzero(link)
        idx = (x + xs.start) + W * ((y + ys.start) + H * (z + zs.start))
        out[i] = (idx.astype(np.int64), q.astype(np.float32))
    return out


def bouzidi_branch_switches(sdf_a, sdf_b, W: int, H: int, D: int = 1) -> int:
    """Number of Bouzidi links of one geometry that the other has with q on the other side of 1/2 or of
    a clamp bound (the interpolation switches formula there, src/shaders/step.ts:146-156, so J(theta)
    has a kink: a central difference between the two geometries straddles a change of slope), plus
    the links only one of them has (a node flipping between fluid and solid)."""
    la, lb = bouzidi_links(sdf_a, W, H, D), bouzidi_links(sdf_b, W, H, D)
    n = 0
    for i in la:
        ia, qa = la[i]
        ib, qb = lb[i]
        common, pa, pb = np.intersect1d(ia, ib, assume_unique=True, return_indices=True)
        n += (len(ia) - len(common)) + (len(ib) - len(common))
        a, b = qa[pa], qb[pb]
        n += int(np.count_nonzero(((a < 0.5) != (b < 0.5)) | ((a <= 1e-3) != (b <= 1e-3)) | ((a >= 1.0) != (b >= 1.0))))
    return n


try:
    import torch

    class TunnelObjective(torch.autograd.Function):
        """J(sdf) with autograd semantics: forward runs lbm_adjoint_forward on the SDF, backward runs
        lbm_adjoint_backward with the incoming gradient as its seed."""

        @staticmethod
        def forward(ctx, sdf, adj: ShapeAdjoint):
            ctx.adj = adj
            ctx.meta = (sdf.shape, sdf.dtype, sdf.device)
            J = adj.forward(sdf.detach().cpu().numpy().astype(np.float32).reshape(-1))
            return torch.tensor(J, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gJ):
            shape, dtype, device = ctx.meta
            g = ctx.adj.backward(float(gJ))
            return torch.from_numpy(g).reshape(shape).to(dtype=dtype, device=device), None

except ImportError:  # torch is optional for the C ABI users
    TunnelObjective = None
