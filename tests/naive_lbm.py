"""Independent naive FP64 lattice Boltzmann, written from the equations (test infrastructure).

Used only to cross-check the oracle on tiny grids. It differs from the oracle in every way that
could hide a shared mistake: numpy arrays of full populations F = f_eq-style (not f - w), a
separate pull-streaming pass per direction with masks, then the boundary rules, then a TRT
collision written as relaxation of symmetric / antisymmetric parts toward the equilibrium, with
the absorbing layer as sigma (F_eq(rho, u_target) - F_eq(rho, u)) and Guo forcing in its
textbook form W_i [3 (c_i - u) . F + 9 (c_i . u)(c_i . F)]. Boundary rules follow the physical
descriptions of the reference (src/shaders/step.ts, src/shaders/step3d.ts): Bouzidi linear
interpolation with q from the SDF, Zou-He velocity inlet and density outlet, specular slip,
halfway bounce-back, moving lid, equilibrium inlet plane, copy-then-equilibrium outlet plane.
Forces use the reference's momentum-exchange definition, c_j ((F_j - w) + (F_i - w)) per link.
"""
from __future__ import annotations

import numpy as np

# ------------------------------------------------------------------ D2Q9
C2 = np.array([[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1], [1, 1], [-1, 1], [-1, -1], [1, -1]])
W2 = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)


def _opp(C):
    return np.array([int(np.where((C == -c).all(axis=1))[0][0]) for c in C])


OPP2 = _opp(C2)
MY2 = np.array([int(np.where((C2 == c * [1, -1]).all(axis=1))[0][0]) for c in C2])


def feq(C, Wt, rho, u):
    """Full equilibrium F_eq[i, ...] for rho[...], u[..., d]."""
    cu = np.tensordot(C.astype(float), np.moveaxis(u, -1, 0), axes=(1, 0))  # (Q, ...)
    usq = (u * u).sum(-1)
    return Wt.reshape((-1,) + (1,) * rho.ndim) * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * usq)


def smoothstep(e0, e1, x):
    t = np.clip((x - e0) / (e1 - e0), 0, 1)
    return t * t * (3 - 2 * t)


def _relax_rates(cfg, W, xs):
    """omega+ per column from the sponge (a ramp of tau toward spongeTau over the last fraction)."""
    tau = cfg["tau"]
    tau_s = max(cfg.get("spongeTau", 1.0), tau)
    frac = cfg.get("spongeFraction", 0.0)
    x0 = (W - 1) * (1 - frac) if frac > 0 else 1e9
    s_out = smoothstep(x0, W - 1, xs) if x0 < W - 1 else np.zeros_like(xs, dtype=float)
    tau_x = tau * (1 - s_out) + tau_s * s_out
    return 1 / tau_x, s_out


def _omega_minus(om_p, lam):
    return 1 / (0.5 + lam / (1 / om_p - 0.5))


def run2d(cfg: dict, sdf, steps: int, macro0=None):
    """Returns (f - w as (9, N), list of (fx, fy) per step)."""
    W, H = cfg["width"], cfg["height"]
    left, right, bottom, top = cfg["left"], cfg["right"], cfg["bottom"], cfg["top"]
    tau, lam = cfg["tau"], cfg.get("lambda", 3 / 16)
    u_in, parabolic = cfg.get("uIn", 0.0), cfg.get("parabolic", False)
    u_lid = cfg.get("uLid", 0.0)
    fxb, fyb = cfg.get("bodyForce", (0.0, 0.0))
    absorb = cfg.get("absorb", 0.0)
    layer = cfg.get("inletLayerFraction", 0.0) * (W - 1)
    smag_c2 = cfg.get("smagorinsky", 0.0) ** 2
    forcing = fxb != 0 or fyb != 0
    sdf = np.asarray(sdf if sdf is not None else np.full(W * H, 1e6), dtype=np.float32).reshape(H, W).astype(float)
    ys, xs = np.mgrid[0:H, 0:W]
    open_col = ((left == 2) & (xs == 0)) | ((right == 2) & (xs == W - 1))
    solid = (sdf < 0) & ~open_col
    fluid = ~solid
    if macro0 is None:
        rho0, u0 = np.ones((H, W)), np.zeros((H, W, 2))
    else:
        m = np.asarray(macro0, dtype=np.float32).reshape(H, W, 4).astype(float)
        rho0, u0 = m[..., 0], m[..., 1:3]
    F = feq(C2, W2, rho0, u0)
    om_p0, s_out = _relax_rates(cfg, W, xs.astype(float))
    s_in = 1 - smoothstep(0, layer, xs.astype(float)) if layer > 0 else np.zeros((H, W))
    sigma = absorb * np.maximum(s_out ** 2, s_in ** 2)
    ut = (4 * u_in * (ys + 0.5) * (H - ys - 0.5) / H ** 2) if parabolic else np.full((H, W), u_in)
    pairs = [(1, 3), (2, 4), (5, 7), (6, 8)]
    forces = []

    def wrap(qx, qy):
        qx = np.where((qx < 0) & (left == 0), qx + W, qx)
        qx = np.where((qx >= W) & (right == 0), qx - W, qx)
        qy = np.where((qy < 0) & (bottom == 0), qy + H, qy)
        qy = np.where((qy >= H) & (top == 0), qy - H, qy)
        return qx, qy

    for _ in range(steps):
        Fin = np.empty_like(F)
        force = np.zeros((H, W, 2))
        Fin[0] = F[0]
        for i in range(1, 9):
            j = OPP2[i]
            sx, sy = wrap(xs - C2[i, 0], ys - C2[i, 1])
            inx = (sx >= 0) & (sx < W)
            iny = (sy >= 0) & (sy < H)
            sxc, syc = np.clip(sx, 0, W - 1), np.clip(sy, 0, H - 1)
            inside = inx & iny
            up_solid = inside & solid[syc, sxc]
            v = F[j].copy()  # bounce-back and open-side placeholders
            v = np.where(inside & ~up_solid, F[i, syc, sxc], v)
            # Bouzidi
            dF, dS = sdf, sdf[syc, sxc]
            q = np.clip(dF / np.maximum(dF - dS, 1e-6), 1e-3, 1.0)
            nx, ny = wrap(xs + C2[i, 0], ys + C2[i, 1])
            nin = (nx >= 0) & (nx < W) & (ny >= 0) & (ny < H)
            nxc, nyc = np.clip(nx, 0, W - 1), np.clip(ny, 0, H - 1)
            n_ok = nin & ~solid[nyc, nxc]
            lo = np.where(n_ok, 2 * q * F[j] + (1 - 2 * q) * F[j, nyc, nxc], F[j])
            hi = F[j] / (2 * q) + (2 * q - 1) / (2 * q) * F[i]
            bz = np.where(q < 0.5, lo, hi)
            v = np.where(up_solid, bz, v)
            link = up_solid & fluid
            for d in range(2):
                force[..., d] += np.where(link, C2[j, d] * ((F[j] - W2[j]) + (bz - W2[i])), 0.0)
            # y sides
            mode = np.where(sy < 0, bottom, top)
            ymask = ~iny
            slip = ymask & (mode == 2)
            v = np.where(slip, F[MY2[i], ys, np.clip(sx, 0, W - 1)], v)
            v = np.where(ymask & (mode == 3), F[j] + 6 * W2[i] * C2[i, 0] * u_lid, v)
            Fin[i] = v
        forces.append(force[fluid].sum(axis=0))
        # Zou-He
        if left == 2:
            c = Fin[:, :, 0]
            u = ut[:, 0]
            rho = (c[0] + c[2] + c[4] + 2 * (c[3] + c[6] + c[7])) / (1 - u)
            c[1] = c[3] + 2 / 3 * rho * u
            c[5] = c[7] - 0.5 * (c[2] - c[4]) + rho * u / 6
            c[8] = c[6] + 0.5 * (c[2] - c[4]) + rho * u / 6
            Fin[:, :, 0] = c
        if right == 2:
            c = Fin[:, :, W - 1]
            j = c[0] + c[2] + c[4] + 2 * (c[1] + c[5] + c[8]) - 1.0
            c[3] = c[1] - 2 / 3 * j
            c[7] = c[5] + 0.5 * (c[2] - c[4]) - j / 6
            c[6] = c[8] - 0.5 * (c[2] - c[4]) - j / 6
            Fin[:, :, W - 1] = c
        rho = Fin.sum(0)
        jv = np.stack([np.tensordot(C2[:, d], Fin, axes=(0, 0)) for d in range(2)], -1)
        u = (jv + 0.5 * np.array([fxb, fyb]) * forcing) / rho[..., None]
        Fe = feq(C2, W2, rho, u)
        om_p = om_p0
        if smag_c2 > 0:
            n = Fin - Fe
            Pi = np.zeros((2, 2) + rho.shape)
            for a in range(2):
                for b in range(2):
                    Pi[a, b] = np.tensordot(C2[:, a] * C2[:, b], n, axes=(0, 0))
            norm = np.sqrt((Pi * Pi).sum(axis=(0, 1)))
            t0 = 1 / om_p0
            om_p = 1 / (0.5 * (t0 + np.sqrt(t0 * t0 + 18 * np.sqrt(2) * smag_c2 * norm / rho)))
        om_m = _omega_minus(om_p, lam)
        out = np.empty_like(Fin)
        out[0] = Fin[0] - om_p * (Fin[0] - Fe[0])
        for a, b in pairs:
            fp, fm = 0.5 * (Fin[a] + Fin[b]), 0.5 * (Fin[a] - Fin[b])
            ep, em = 0.5 * (Fe[a] + Fe[b]), 0.5 * (Fe[a] - Fe[b])
            out[a] = Fin[a] - om_p * (fp - ep) - om_m * (fm - em)
            out[b] = Fin[b] - om_p * (fp - ep) + om_m * (fm - em)
        # absorbing layer toward the inflow state
        ut2 = np.stack([ut, np.zeros_like(ut)], -1)
        out += sigma * (feq(C2, W2, rho, ut2) - Fe)
        if forcing:
            Fv = np.array([fxb, fyb])
            cu = np.tensordot(C2.astype(float), np.moveaxis(u, -1, 0), axes=(1, 0))
            cF = C2 @ Fv
            uF = (u * Fv).sum(-1)
            sym = W2[:, None, None] * (-3 * uF + 9 * cu * cF[:, None, None])
            anti = W2[:, None, None] * 3 * cF[:, None, None] * np.ones_like(cu)
            for a, b in [(0, 0)] + pairs:
                sp = 0.5 * (sym[a] + sym[b])
                sa = 0.5 * (anti[a] - anti[b])
                out[a] += (1 - 0.5 * om_p) * sp + (1 - 0.5 * om_m) * sa
                if b != a:
                    out[b] += (1 - 0.5 * om_p) * sp - (1 - 0.5 * om_m) * sa
        F = np.where(fluid[None], out, F)
    return (F - W2[:, None, None]).reshape(9, -1), np.array(forces)


# ------------------------------------------------------------------ D3Q19
C3 = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [1, 1, 0],
               [-1, -1, 0], [1, 0, 1], [-1, 0, -1], [0, 1, 1], [0, -1, -1], [1, -1, 0], [-1, 1, 0], [1, 0, -1],
               [-1, 0, 1], [0, 1, -1], [0, -1, 1]])
W3 = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
OPP3 = _opp(C3)
MY3 = np.array([int(np.where((C3 == c * [1, -1, 1]).all(axis=1))[0][0]) for c in C3])
MZ3 = np.array([int(np.where((C3 == c * [1, 1, -1]).all(axis=1))[0][0]) for c in C3])


# D3Q27, built independently of any table: every velocity of {-1, 0, 1}^3, weighted by how many
# components are non-zero (8/27, 2/27, 1/54, 1/216). Its order is itertools.product's, not the oracle's.
C27 = np.array([[cx, cy, cz] for cx in (-1, 0, 1) for cy in (-1, 0, 1) for cz in (-1, 0, 1)])
W27 = np.array([(8 / 27, 2 / 27, 1 / 54, 1 / 216)[int(np.abs(c).sum())] for c in C27])


def run3d(cfg: dict, sdf, steps: int, C3=C3, W3=W3):
    """Returns (f - w as (Q, N), list of (fx, fy, fz) per step) for the velocity set C3, W3."""
    Q = len(C3)
    OPP3 = _opp(C3)
    MY3 = np.array([int(np.where((C3 == c * [1, -1, 1]).all(axis=1))[0][0]) for c in C3])
    MZ3 = np.array([int(np.where((C3 == c * [1, 1, -1]).all(axis=1))[0][0]) for c in C3])
    W, H, D = cfg["width"], cfg["height"], cfg["depth"]
    tau = cfg["tau"]
    lam = cfg.get("lambda")
    if lam is None or lam != lam:
        lam = min(3 / 16, 50 * (tau - 0.5) ** 2)
    u_in = cfg.get("uIn", 0.0)
    absorb = cfg.get("absorb", 0.0)
    smag_c2 = cfg.get("smagorinsky", 0.0) ** 2
    sdf = np.asarray(sdf, dtype=np.float32).reshape(D, H, W).astype(float)
    zs, ys, xs = np.mgrid[0:D, 0:H, 0:W]
    solid = (sdf < 0) & (xs != 0) & (xs != W - 1)
    fluid = ~solid
    inlet = xs == 0
    outlet = xs == W - 1
    rho0 = np.ones((D, H, W))
    u0 = np.zeros((D, H, W, 3))
    u0[..., 0] = u_in
    F = np.where(fluid[None], feq(C3, W3, rho0, u0), W3[:, None, None, None])  # solids hold f - w = 0
    om_p0, s_out = _relax_rates(cfg, W, xs.astype(float))
    sigma = absorb * s_out ** 2
    rest = int(np.where((C3 == 0).all(axis=1))[0][0])
    pairs = [(a, int(OPP3[a])) for a in range(Q) if a < OPP3[a]]
    forces = []
    for _ in range(steps):
        Fin = np.empty_like(F)
        force = np.zeros((D, H, W, 3))
        Fin[rest] = F[rest]
        for i in range(Q):
            if i == rest:
                continue
            j = OPP3[i]
            sx, sy, sz = xs - C3[i, 0], ys - C3[i, 1], zs - C3[i, 2]
            ox = (sx < 0) | (sx >= W)
            oy = (sy < 0) | (sy >= H)
            oz = (sz < 0) | (sz >= D)
            sxc, syc, szc = np.clip(sx, 0, W - 1), np.clip(sy, 0, H - 1), np.clip(sz, 0, D - 1)
            v = F[j].copy()  # bounce-back (edges, x walls)
            # outlet copy from the plane behind
            v = np.where(ox & outlet, F[i, zs, ys, W - 2 if W >= 2 else 0], v)
            # slip faces (one of y/z out, x in): mirror from the reflected node
            face = ~ox & (oy ^ oz)
            rx, ry, rz = sxc, np.where(oy, ys, syc), np.where(oz, zs, szc)
            r_solid = solid[rz, ry, rx]
            mdir = np.where(oy, MY3[i], MZ3[i])
            mirrored = np.where(oy, F[MY3[i], rz, ry, rx], F[MZ3[i], rz, ry, rx])
            del mdir
            fs = face & ~(ox & outlet)
            v = np.where(fs & ~r_solid, mirrored, v)
            through = fs & r_solid & fluid
            # slip edges (y and z out, x in): specular reflection through both faces
            edge = ~ox & oy & oz
            e_solid = solid[zs, ys, sxc]
            v = np.where(edge & ~e_solid, F[MY3[MZ3[i]], zs, ys, sxc], v)
            through |= edge & e_solid & fluid
            for d in range(3):
                force[..., d] += np.where(through, C3[j, d] * 2 * (F[j] - W3[j]), 0.0)
            inside = ~ox & ~oy & ~oz
            up_solid = inside & solid[szc, syc, sxc]
            v = np.where(inside & ~up_solid, F[i, szc, syc, sxc], v)
            dF, dS = sdf, sdf[szc, syc, sxc]
            q = np.clip(dF / np.maximum(dF - dS, 1e-6), 1e-3, 1.0)
            nx, ny, nz = xs + C3[i, 0], ys + C3[i, 1], zs + C3[i, 2]
            nin = (nx >= 0) & (nx < W) & (ny >= 0) & (ny < H) & (nz >= 0) & (nz < D)
            nxc, nyc, nzc = np.clip(nx, 0, W - 1), np.clip(ny, 0, H - 1), np.clip(nz, 0, D - 1)
            n_ok = nin & ~solid[nzc, nyc, nxc]
            lo = np.where(n_ok, 2 * q * F[j] + (1 - 2 * q) * F[j, nzc, nyc, nxc], F[j])
            hi = F[j] / (2 * q) + (2 * q - 1) / (2 * q) * F[i]
            bz = np.where(q < 0.5, lo, hi)
            v = np.where(up_solid, bz, v)
            link = up_solid & fluid & ~inlet
            for d in range(3):
                force[..., d] += np.where(link, C3[j, d] * ((F[j] - W3[j]) + (bz - W3[i])), 0.0)
            Fin[i] = v
        forces.append(force[fluid & ~inlet].sum(axis=0))
        rho = Fin.sum(0)
        u = np.stack([np.tensordot(C3[:, d], Fin, axes=(0, 0)) for d in range(3)], -1) / rho[..., None]
        Fe = feq(C3, W3, rho, u)
        om_p = om_p0
        if smag_c2 > 0:
            n = Fin - Fe
            Pi = np.zeros((3, 3) + rho.shape)
            for a in range(3):
                for b in range(3):
                    Pi[a, b] = np.tensordot(C3[:, a] * C3[:, b], n, axes=(0, 0))
            norm = np.sqrt((Pi * Pi).sum(axis=(0, 1)))
            t0 = 1 / om_p0
            om_p = 1 / (0.5 * (t0 + np.sqrt(t0 * t0 + 18 * np.sqrt(2) * smag_c2 * norm / rho)))
        om_m = _omega_minus(om_p, lam)
        out = np.empty_like(Fin)
        out[rest] = Fin[rest] - om_p * (Fin[rest] - Fe[rest])
        for a, b in pairs:
            fp, fm = 0.5 * (Fin[a] + Fin[b]), 0.5 * (Fin[a] - Fin[b])
            ep, em = 0.5 * (Fe[a] + Fe[b]), 0.5 * (Fe[a] - Fe[b])
            out[a] = Fin[a] - om_p * (fp - ep) - om_m * (fm - em)
            out[b] = Fin[b] - om_p * (fp - ep) + om_m * (fm - em)
        ut = np.zeros_like(u)
        ut[..., 0] = u_in
        out += sigma * (feq(C3, W3, rho, ut) - Fe)
        # outlet plane: equilibrium at rho = 1 with the local velocity; inlet plane: inflow equilibrium
        out = np.where(outlet[None], feq(C3, W3, np.ones_like(rho), u), out)
        u_inflow = np.zeros_like(u)
        u_inflow[..., 0] = u_in
        out = np.where(inlet[None], feq(C3, W3, np.ones_like(rho), u_inflow), out)
        F = np.where(fluid[None], out, F)
    return (F - W3[:, None, None, None]).reshape(Q, -1), np.array(forces)
