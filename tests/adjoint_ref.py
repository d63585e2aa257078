"""FP64 CPU reference gradient for the D2Q9 shape adjoint (test infrastructure).

A differentiable torch (float64, CPU) restatement of the step of tests/naive_lbm.py for the
tunnel setting the adjoint supports (Zou-He inlet/outlet, slip or no-slip y walls, sponge and
absorbing layers, Bouzidi walls, no Smagorinsky, no body force), with the SDF as an autograd leaf:
J = mean over a step window of C_D = 2 F_x / (u^2 L) from the momentum-exchange force. Reverse
mode through torch.utils.checkpoint segments gives dJ/dsdf at every node, the FP64 reference the
GPU's hand-written discrete adjoint is checked against. The flag topology (which nodes are solid)
is fixed from the SDF's signs, as the solver fixes it at set_sdf.
"""
from __future__ import annotations

import numpy as np
import torch
from torch.utils.checkpoint import checkpoint

from tests import naive_lbm as nv

C2 = torch.tensor(nv.C2, dtype=torch.float64)
W2 = torch.tensor(nv.W2, dtype=torch.float64)


def _feq(rho, ux, uy):
    cu = C2[:, 0, None, None] * ux + C2[:, 1, None, None] * uy
    usq = ux * ux + uy * uy
    return W2[:, None, None] * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * usq)


class Tunnel2D:
    def __init__(self, cfg: dict, sdf: np.ndarray):
        self.cfg = cfg
        W, H = cfg["width"], cfg["height"]
        self.W, self.H = W, H
        left, right, bottom, top = cfg["left"], cfg["right"], cfg["bottom"], cfg["top"]
        assert left == 2 and right == 2 and bottom != 0 and top != 0, "open x, walls in y"
        assert cfg.get("smagorinsky", 0.0) == 0 and tuple(cfg.get("bodyForce", (0, 0))) == (0, 0)
        sd = np.asarray(sdf, dtype=np.float32).reshape(H, W).astype(np.float64)
        ys, xs = np.mgrid[0:H, 0:W]
        open_col = (xs == 0) | (xs == W - 1)
        self.solid = torch.tensor((sd < 0) & ~open_col)
        self.fluid = ~self.solid
        om_p0, s_out = nv._relax_rates(cfg, W, xs.astype(float))
        self.om_p = torch.tensor(om_p0)
        lam = cfg.get("lambda", 3 / 16)
        self.om_m = 1 / (0.5 + lam / (1 / self.om_p - 0.5))
        s_in = np.zeros((H, W))
        layer = cfg.get("inletLayerFraction", 0.0) * (W - 1)
        if layer > 0:
            s_in = 1 - nv.smoothstep(0, layer, xs.astype(float))
        self.sigma = torch.tensor(cfg.get("absorb", 0.0) * np.maximum(s_out ** 2, s_in ** 2))
        self.parabolic = cfg.get("parabolic", False)
        self.ys = torch.tensor(ys, dtype=torch.float64)
        self.xs_i, self.ys_i = xs, ys
        # per direction: upstream index arrays and masks (no periodic wrap: x open, y walls)
        self.dirs = []
        solid = self.solid.numpy()
        for i in range(1, 9):
            j = int(nv.OPP2[i])
            sx, sy = xs - nv.C2[i, 0], ys - nv.C2[i, 1]
            inx = (sx >= 0) & (sx < W)
            iny = (sy >= 0) & (sy < H)
            sxc, syc = np.clip(sx, 0, W - 1), np.clip(sy, 0, H - 1)
            inside = inx & iny
            up_solid = inside & solid[syc, sxc]
            nx, ny = xs + nv.C2[i, 0], ys + nv.C2[i, 1]
            nin = (nx >= 0) & (nx < W) & (ny >= 0) & (ny < H)
            nxc, nyc = np.clip(nx, 0, W - 1), np.clip(ny, 0, H - 1)
            n_ok = nin & ~solid[nyc, nxc]
            mode = np.where(sy < 0, bottom, top)
            slip = ~iny & (mode == 2)
            moving = ~iny & (mode == 3)
            self.dirs.append(dict(i=i, j=j, syc=syc, sxc=sxc, direct=torch.tensor(inside & ~up_solid),
                                  bz=torch.tensor(up_solid), nyc=nyc, nxc=nxc, n_ok=torch.tensor(n_ok),
                                  slip=torch.tensor(slip), moving=torch.tensor(moving),
                                  slip_x=np.clip(sx, 0, W - 1)))

    def inlet_u(self, u_in):
        if self.parabolic:
            H = self.H
            s = self.ys[:, 0] + 0.5
            return 4 * u_in * s * (H - s) / (H * H)
        return torch.full((self.H,), u_in, dtype=torch.float64)

    def step(self, F, sdf, u_in):
        """One step on full populations F[9, H, W]; returns (F_next, F_x, F_y)."""
        H, W = self.H, self.W
        Fin = [F[0]]
        fx = torch.zeros((), dtype=torch.float64)
        fy = torch.zeros((), dtype=torch.float64)
        for d in self.dirs:
            i, j = d["i"], d["j"]
            v = F[j].clone()
            v = torch.where(d["direct"], F[i][d["syc"], d["sxc"]], v)
            dF = sdf
            dS = sdf[d["syc"], d["sxc"]]
            q = torch.clamp(dF / torch.clamp(dF - dS, min=1e-6), 1e-3, 1.0)
            lo = torch.where(d["n_ok"], 2 * q * F[j] + (1 - 2 * q) * F[j][d["nyc"], d["nxc"]], F[j])
            hi = F[j] / (2 * q) + (2 * q - 1) / (2 * q) * F[i]
            bz = torch.where(q < 0.5, lo, hi)
            v = torch.where(d["bz"], bz, v)
            link = d["bz"] & self.fluid
            contrib = torch.where(link, (F[j] - W2[j]) + (bz - W2[i]), torch.zeros_like(bz))
            fx = fx + float(nv.C2[j, 0]) * contrib.sum()
            fy = fy + float(nv.C2[j, 1]) * contrib.sum()
            v = torch.where(d["slip"], F[nv.MY2[i]][self.ys_i, d["slip_x"]], v)
            v = torch.where(d["moving"], F[j] + 6 * W2[i] * float(nv.C2[i, 0]) * self.cfg.get("uLid", 0.0), v)
            Fin.append(v)
        Fin = torch.stack(Fin)
        # Zou-He
        u = self.inlet_u(u_in)
        c = Fin[:, :, 0]
        rho_in = (c[0] + c[2] + c[4] + 2 * (c[3] + c[6] + c[7])) / (1 - u)
        c1 = c[3] + 2 / 3 * rho_in * u
        c5 = c[7] - 0.5 * (c[2] - c[4]) + rho_in * u / 6
        c8 = c[6] + 0.5 * (c[2] - c[4]) + rho_in * u / 6
        col0 = torch.stack([c[0], c1, c[2], c[3], c[4], c5, c[6], c[7], c8])
        c = Fin[:, :, W - 1]
        jo = c[0] + c[2] + c[4] + 2 * (c[1] + c[5] + c[8]) - 1.0
        c3 = c[1] - 2 / 3 * jo
        c7 = c[5] + 0.5 * (c[2] - c[4]) - jo / 6
        c6 = c[8] - 0.5 * (c[2] - c[4]) - jo / 6
        colW = torch.stack([c[0], c[1], c[2], c3, c[4], c[5], c6, c7, c[8]])
        Fin = torch.cat([col0[:, :, None], Fin[:, :, 1 : W - 1], colW[:, :, None]], dim=2)
        rho = Fin.sum(0)
        ux = (C2[:, 0, None, None] * Fin).sum(0) / rho
        uy = (C2[:, 1, None, None] * Fin).sum(0) / rho
        Fe = _feq(rho, ux, uy)
        out = [Fin[0] - self.om_p * (Fin[0] - Fe[0])]
        pairs = [(1, 3), (2, 4), (5, 7), (6, 8)]
        res = [None] * 9
        res[0] = out[0]
        for a, b in pairs:
            fp, fm = 0.5 * (Fin[a] + Fin[b]), 0.5 * (Fin[a] - Fin[b])
            ep, em = 0.5 * (Fe[a] + Fe[b]), 0.5 * (Fe[a] - Fe[b])
            res[a] = Fin[a] - self.om_p * (fp - ep) - self.om_m * (fm - em)
            res[b] = Fin[b] - self.om_p * (fp - ep) + self.om_m * (fm - em)
        out = torch.stack(res)
        ut = u[:, None].expand(H, W)
        out = out + self.sigma * (_feq(rho, ut, torch.zeros_like(ut)) - Fe)
        F_next = torch.where(self.fluid[None], out, F)
        return F_next, fx, fy

    def objective(self, sdf, steps, window, u, ramp_steps, L, segment=100, ramp_kind="smoothstep"):
        """J = mean of 2 F_x / (u^2 L) over the samples of steps (steps - window, steps]; inlet ramp
        u r(min(1, k / ramp_steps)) with r(t) = t^2 (3 - 2t) (smoothstep) or t (linear)."""
        H, W = self.H, self.W
        F = _feq(torch.ones(H, W, dtype=torch.float64), torch.zeros(H, W, dtype=torch.float64),
                 torch.zeros(H, W, dtype=torch.float64))

        def seg(F, sdf, k0, k1):
            acc = torch.zeros((), dtype=torch.float64)
            for k in range(k0, k1):
                r = min(1.0, k / ramp_steps) if ramp_steps > 0 else 1.0
                uin = u * r * r * (3 - 2 * r) if ramp_kind == "smoothstep" else u * r
                F, fx, _ = self.step(F, sdf, uin)
                if k + 1 > steps - window:
                    acc = acc + fx
            return F, acc

        total = torch.zeros((), dtype=torch.float64)
        for k0 in range(0, steps, segment):
            k1 = min(steps, k0 + segment)
            F, acc = checkpoint(seg, F, sdf, k0, k1, use_reentrant=False)
            total = total + acc
        return 2 * total / window / (u * u * L), F


def reference_gradient(cfg, sdf, steps, window, u, ramp_steps, L, segment=20, ramp_kind="smoothstep"):
    """(J, dJ/dsdf [N]) in float64 on the CPU: a forward pass without a graph that keeps the state at
    the start of every segment, then per segment in reverse a re-run with autograd and the
    vector-Jacobian product of (its share of J + gF . F_end) for (F_start, sdf)."""
    t = Tunnel2D(cfg, sdf)
    H, W = t.H, t.W
    s = torch.tensor(np.asarray(sdf, dtype=np.float32).astype(np.float64).reshape(H, W), requires_grad=True)
    scale = 2.0 / window / (u * u * L)

    def u_in(k):
        r = min(1.0, k / ramp_steps) if ramp_steps > 0 else 1.0
        return u * r * r * (3 - 2 * r) if ramp_kind == "smoothstep" else u * r

    F = _feq(torch.ones(H, W, dtype=torch.float64), torch.zeros(H, W, dtype=torch.float64),
             torch.zeros(H, W, dtype=torch.float64))
    starts, total = [], 0.0
    with torch.no_grad():
        for k0 in range(0, steps, segment):
            starts.append(F.clone())
            for k in range(k0, min(steps, k0 + segment)):
                F, fx, _ = t.step(F, s, u_in(k))
                if k + 1 > steps - window:
                    total += float(fx)
    gF = torch.zeros_like(F)
    g_sdf = torch.zeros_like(s)
    for i in reversed(range(len(starts))):
        k0 = i * segment
        Fs = starts[i].clone().requires_grad_(True)
        F = Fs
        acc = torch.zeros((), dtype=torch.float64)
        for k in range(k0, min(steps, k0 + segment)):
            F, fx, _ = t.step(F, s, u_in(k))
            if k + 1 > steps - window:
                acc = acc + fx
        gFs, gs = torch.autograd.grad(scale * acc + (gF * F).sum(), [Fs, s], allow_unused=True)
        gF = gFs
        if gs is not None:
            g_sdf += gs
    return scale * total, g_sdf.numpy().reshape(-1).copy()
