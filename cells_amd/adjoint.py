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
