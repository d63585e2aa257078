'use strict';
/**
 * Validation suite of the reference (src/cases.ts:61-427, src/cases3d.ts:11-81) on the gfx950 solver,
 * driven from Node through the same Solver / Solver3D classes the reference's pages use. Every case
 * keeps the reference's parameters, inlet ramp, convergence loop, metric definitions and acceptance
 * ranges (fixed before a case runs, AGENTS.md:90). A case resolves to
 * { name, metrics: [{ name, value, ref, accept, pass }], pass, steps, cells, seconds, notes }.
 *
 *   node tests/validation/validate.js [case ...]     (default: every case; one JSON line per case)
 */
const { Solver, Solver3D, XMode, YMode } = require('../../cells_amd/node/solver');
const g = require('../../cells_amd/node/geometry');

/** metric (src/cases.ts:31-33). */
function metric(name, value, ref, accept) {
  return { name, value, ref, accept, pass: accept[0] <= value && value <= accept[1] };
}
const around = (center, rel) => [center * (1 - rel), center * (1 + rel)];

/** runRamped (src/cases.ts:41-59): smoothstep inlet ramp updated per chunk, forces collected. */
async function runRamped(s, steps, uTarget, rampSteps, forces, chunk) {
  const c = chunk || 400;
  const end = s.step + steps;
  while (s.step < end) {
    const r = Math.min(1, s.step / rampSteps);
    const uIn = uTarget * r * r * (3 - 2 * r);
    if (uIn !== s.cfg.uIn) s.update({ uIn });
    await s.run(Math.min(c, end - s.step), c);
    for (const f of await s.readForces()) forces.push(f);
  }
  return forces;
}

/** src/cases.ts:61-92. */
async function poiseuille() {
  const H = 32;
  const uMax = 0.05;
  const metrics = [];
  let steps = 0;
  for (const tau of [0.6, 1.0, 1.5]) {
    const nu = (tau - 0.5) / 3;
    const fx = (8 * nu * uMax) / (H * H);
    const s = await Solver.create(0, { width: 4, height: H, left: XMode.Periodic, right: XMode.Periodic,
      bottom: YMode.NoSlip, top: YMode.NoSlip, tau, bodyForce: [fx, 0], forceEvery: 0 });
    s.initField();
    const n = Math.ceil((10 * H * H) / nu);
    await s.run(n, 2000);
    steps += n;
    const m = await s.readMacro();
    let num = 0;
    let den = 0;
    for (let y = 0; y < H; y++) {
      const e = (fx / (2 * nu)) * (y + 0.5) * (H - y - 0.5);
      const ux = m[(y * 4 + 1) * 4 + 1];
      num += (ux - e) * (ux - e);
      den += e * e;
    }
    metrics.push(metric(`L2 error, tau=${tau}`, Math.sqrt(num / den), [0, 0], [0, 2e-4]));
    s.destroy();
  }
  return { name: 'Plane Poiseuille', metrics, steps, cells: 4 * H, notes: [] };
}

/** src/cases.ts:94-126. */
async function periodicSeam() {
  const [W, H, D, tau] = [128, 64, 16, 0.8];
  const drag = [];
  let steps = 0;
  for (const cx of [W / 2 + D / 2 - 0.5, D / 2 - 0.5]) {
    const s = await Solver.create(0, { width: W, height: H, left: XMode.Periodic, right: XMode.Periodic,
      bottom: YMode.NoSlip, top: YMode.NoSlip, tau, bodyForce: [2e-6, 0], forceEvery: 100 });
    const sdf = g.emptySdf(W, H);
    g.addCircle(sdf, W, H, cx, H / 2 - 0.5, D / 2);
    g.addCircle(sdf, W, H, cx + W, H / 2 - 0.5, D / 2);
    s.setSdf(sdf);
    s.initField();
    const forces = [];
    for (let k = 0; k < 20; k++) {
      await s.run(1000, 1000);
      for (const f of await s.readForces()) forces.push(f);
    }
    steps += s.step;
    drag.push(g.mean(forces.slice(-5).map((f) => f.fx)));
    s.destroy();
  }
  return { name: 'Periodic seam invariance',
    metrics: [metric('drag at seam / drag in middle', drag[1] / drag[0], [1, 1], [0.999, 1.001])],
    steps, cells: W * H, notes: [`mid-domain drag ${drag[0].toExponential(7)}, seam drag ${drag[1].toExponential(7)}`] };
}

/** src/cases.ts:128-192. */
async function taylorGreen() {
  const tau = 0.8;
  const nu = (tau - 0.5) / 3;
  const errors = [];
  const nuMeasured = [];
  let steps = 0;
  for (const n of [32, 64, 128]) {
    const u0 = (0.04 * 32) / n;
    const k = (2 * Math.PI) / n;
    const s = await Solver.create(0, { width: n, height: n, left: XMode.Periodic, right: XMode.Periodic,
      bottom: YMode.Periodic, top: YMode.Periodic, tau, forceEvery: 0 });
    const field = (t) => {
      const a = u0 * Math.exp(-2 * nu * k * k * t);
      const m = new Float32Array(n * n * 4);
      for (let y = 0; y < n; y++) {
        for (let x = 0; x < n; x++) {
          const i = (y * n + x) * 4;
          m[i] = 1 - 0.75 * a * a * (Math.cos(2 * k * x) + Math.cos(2 * k * y));
          m[i + 1] = -a * Math.cos(k * x) * Math.sin(k * y);
          m[i + 2] = a * Math.sin(k * x) * Math.cos(k * y);
        }
      }
      return m;
    };
    const energy = (m) => {
      let e = 0;
      for (let i = 0; i < n * n; i++) e += m[i * 4 + 1] * m[i * 4 + 1] + m[i * 4 + 2] * m[i * 4 + 2];
      return e;
    };
    s.initField(field(0));
    const t = Math.round(Math.log(2) / (2 * nu * k * k));
    await s.run(Math.round(t / 2), 1000);
    const mid = await s.readMacro();
    await s.run(t - Math.round(t / 2), 1000);
    const end = await s.readMacro();
    steps += t;
    const exact = field(t);
    let num = 0;
    let den = 0;
    for (let i = 0; i < n * n; i++) {
      for (const c of [1, 2]) {
        num += (end[i * 4 + c] - exact[i * 4 + c]) ** 2;
        den += exact[i * 4 + c] ** 2;
      }
    }
    errors.push(Math.sqrt(num / den));
    const dt = t - Math.round(t / 2);
    nuMeasured.push(Math.log(energy(mid) / energy(end)) / (4 * k * k * dt));
    s.destroy();
  }
  const order = Math.log2(errors[0] / errors[2]) / 2;
  return { name: 'Taylor-Green vortex',
    metrics: [metric('nu measured / nu (N=128)', nuMeasured[2] / nu, [1, 1], [0.99, 1.01]),
      metric('velocity L2 error (N=128)', errors[2], [0, 0], [0, 0.01]),
      metric('convergence order 32 to 128', order, [2, 2], [1.8, Infinity])],
    steps, cells: 128 * 128, notes: [`L2 errors for N=32/64/128: ${errors.map((v) => v.toExponential(2)).join(', ')}`] };
}

/** schaferTurek (src/cases.ts:195-214). */
async function schaferTurek(D, re, forceEvery, uMax) {
  const W = 22 * D;
  const H = Math.round(4.1 * D);
  const uMean = (2 / 3) * uMax;
  const nu = (uMean * D) / re;
  const cx = 2 * D;
  const cy = 2 * D - 0.5;
  const s = await Solver.create(0, { width: W, height: H, left: XMode.Open, right: XMode.Open, bottom: YMode.NoSlip,
    top: YMode.NoSlip, tau: 3 * nu + 0.5, parabolic: true, uIn: 0, forceEvery, spongeFraction: 0.1, spongeTau: 1,
    absorb: 0.1 });
  const sdf = g.emptySdf(W, H);
  g.addCircle(sdf, W, H, cx, cy, D / 2);
  s.setSdf(sdf);
  s.initField();
  return { s, W, H, cx, cy, uMean, coef: (f) => (2 * f) / (uMean * uMean * D) };
}

/** src/cases.ts:216-269. */
async function schaferTurek1() {
  const D = 40;
  const { s, W, H, cx, cy, uMean, coef } = await schaferTurek(D, 20, 50, 0.05);
  const forces = [];
  let prev = Infinity;
  for (let k = 0; k < 20; k++) {
    await runRamped(s, 10000, 0.05, 5000, forces);
    const cd = coef(g.mean(forces.slice(-20).map((f) => f.fx)));
    if (k >= 4 && Math.abs(cd - prev) < 1e-4) break;
    prev = cd;
  }
  const last = forces.slice(-20);
  const cd = coef(g.mean(last.map((f) => f.fx)));
  const cl = coef(g.mean(last.map((f) => f.fy)));
  const m = await s.readMacro();
  const y0 = Math.floor(cy);
  const w = cy - y0;
  const rhoAt = (x) => (1 - w) * m[(y0 * W + x) * 4] + w * m[((y0 + 1) * W + x) * 4];
  const uxAt = (x) => (1 - w) * m[(y0 * W + x) * 4 + 1] + w * m[((y0 + 1) * W + x) * 4 + 1];
  const xf = cx - D / 2;
  const xb = cx + D / 2;
  const a = Math.floor(xf);
  const b = Math.ceil(xb);
  const pFront = rhoAt(a) + (rhoAt(a) - rhoAt(a - 1)) * (xf - a);
  const pBack = rhoAt(b) + (rhoAt(b) - rhoAt(b + 1)) * (b - xb);
  const dp = (pFront - pBack) / 3 / (uMean * uMean);
  let la = NaN;
  for (let x = b; x < W - 1; x++) {
    const u0 = uxAt(x);
    const u1 = uxAt(x + 1);
    if (u0 < 0 && u1 >= 0) {
      la = (x + -u0 / (u1 - u0) - xb) / D;
      break;
    }
  }
  const steps = s.step;
  s.destroy();
  return { name: 'Schäfer-Turek 2D-1 (Re 20)',
    metrics: [metric('c_D', cd, [5.57, 5.59], around(5.58, 0.02)),
      metric('c_L', cl, [0.0104, 0.011], around(0.0107, 0.1)),
      metric('L_a / D', la, [0.842, 0.852], around(0.847, 0.03)),
      metric('dP / (rho U_mean^2)', dp, [0.1172 / 0.04, 0.1176 / 0.04], around(0.1174 / 0.04, 0.02))],
    steps, cells: W * H, notes: [`D = ${D} cells, grid ${W}x${H}`] };
}

/** src/cases.ts:271-306. */
async function schaferTurek2() {
  const D = 40;
  const { s, W, H, uMean, coef } = await schaferTurek(D, 100, 5, 0.05);
  const period = D / (0.3 * uMean);
  let forces = [];
  let window = [];
  let prevAmp = 0;
  for (let k = 0; k < 40; k++) {
    await runRamped(s, Math.round(5 * period), 0.05, 5000, forces);
    window = forces.filter((f) => f.step > s.step - 5 * period);
    const cls = window.map((f) => coef(f.fy));
    const amp = Math.max(...cls) - Math.min(...cls);
    if (k >= 6 && amp > 0.5 && Math.abs(amp - prevAmp) < 1e-3 * amp) break;
    prevAmp = amp;
    forces = forces.slice(Math.max(0, forces.length - 20000));
  }
  const t = window.map((f) => f.step);
  const cl = window.map((f) => coef(f.fy));
  const cd = window.map((f) => coef(f.fx));
  const p = g.periodFromCrossings(t, cl);
  const st = p ? D / (p.period * uMean) : NaN;
  const steps = s.step;
  s.destroy();
  return { name: 'Schäfer-Turek 2D-2 (Re 100)',
    metrics: [metric('c_D,max', Math.max(...cd), [3.22, 3.24], around(3.23, 0.02)),
      metric('c_L,max', Math.max(...cl), [0.99, 1.01], around(1.0, 0.05)),
      metric('St', st, [0.295, 0.305], around(0.3, 0.02))],
    steps, cells: W * H, notes: [`D = ${D} cells, grid ${W}x${H}, forces every 5 steps`] };
}

/** src/cases.ts:308-355. */
async function unconfinedCylinder() {
  const [D, lx, ly, xc, u] = [32, 40, 48, 16, 0.05];
  const W = lx * D;
  const H = ly * D;
  const nu = (u * D) / 100;
  const s = await Solver.create(0, { width: W, height: H, left: XMode.Open, right: XMode.Open, bottom: YMode.Slip,
    top: YMode.Slip, tau: 3 * nu + 0.5, forceEvery: 5, spongeFraction: 0.15, spongeTau: 1, absorb: 0.02 });
  const sdf = g.emptySdf(W, H);
  g.addCircle(sdf, W, H, xc * D, H / 2 + 0.3, D / 2);
  s.setSdf(sdf);
  s.initField();
  const coef = (f) => (2 * f) / (u * u * D);
  const period = D / (0.165 * u);
  let forces = [];
  let window = [];
  let prevAmp = 0;
  for (let k = 0; k < 60; k++) {
    await runRamped(s, Math.round(5 * period), u, 5000, forces);
    window = forces.filter((f) => f.step > s.step - 5 * period);
    const cls = window.map((f) => coef(f.fy));
    const amp = Math.max(...cls) - Math.min(...cls);
    if (k >= 6 && amp > 0.3 && Math.abs(amp - prevAmp) < 2e-3 * amp) break;
    prevAmp = amp;
    forces = forces.slice(Math.max(0, forces.length - 20000));
  }
  const t = window.map((f) => f.step);
  const p = g.periodFromCrossings(t, window.map((f) => f.fy));
  const st = p ? D / (p.period * u) : NaN;
  const n = p ? Math.floor((t[t.length - 1] - t[0]) / p.period) : 0;
  const span = window.filter((f) => f.step > t[t.length - 1] - n * (p ? p.period : 0));
  const cd = coef(g.mean(span.map((f) => f.fx)));
  const steps = s.step;
  s.destroy();
  return { name: 'Unconfined cylinder (Re 100)',
    metrics: [metric('St', st, [0.1643, 0.167], around(0.165, 0.04)), metric('mean C_D', cd, [1.33, 1.36], around(1.345, 0.05))],
    steps, cells: W * H, notes: [`D = ${D} cells, domain ${lx}D x ${ly}D, slip side walls`] };
}

/** src/cases.ts:357-390. */
async function cavity() {
  const [N, u] = [128, 0.1];
  const nu = (u * N) / 1000;
  const s = await Solver.create(0, { width: N, height: N, left: XMode.Wall, right: XMode.Wall, bottom: YMode.NoSlip,
    top: YMode.MovingWall, tau: 3 * nu + 0.5, uLid: u, forceEvery: 0 });
  s.initField();
  let v = [0, 0, 0];
  let prev = [-1, -1];
  for (let k = 0; k < 40; k++) {
    await s.run(20000, 2000);
    const m = await s.readMacro();
    const ux = new Float64Array(N * N);
    const uy = new Float64Array(N * N);
    for (let i = 0; i < N * N; i++) {
      ux[i] = m[i * 4 + 1];
      uy[i] = m[i * 4 + 2];
    }
    v = g.primaryVortex(g.streamFunction(ux, uy, N, N), N, N);
    if (Math.hypot(v[0] - prev[0], v[1] - prev[1]) < 0.01 && k >= 5) break;
    prev = v;
  }
  const steps = s.step;
  s.destroy();
  return { name: 'Ghia lid-driven cavity (Re 1000)',
    metrics: [metric('vortex x / L', (v[0] + 0.5) / N, [0.5313, 0.5313], [0.5213, 0.5413]),
      metric('vortex y / L', (v[1] + 0.5) / N, [0.5625, 0.5625], [0.5525, 0.5725])],
    steps, cells: N * N, notes: [`N = ${N}, lid u = ${u}`] };
}

/** src/cases.ts:392-425. */
async function naca() {
  const [c, u] = [100, 0.1];
  const W = 16 * c;
  const H = 8 * c;
  const nu = (u * c) / 500;
  const s = await Solver.create(0, { width: W, height: H, left: XMode.Open, right: XMode.Open, bottom: YMode.Slip,
    top: YMode.Slip, tau: 3 * nu + 0.5, forceEvery: 50, spongeFraction: 0.15, spongeTau: 1, absorb: 0.1,
    inletLayerFraction: 0.08 });
  const sdf = g.emptySdf(W, H);
  g.addPolygon(sdf, W, H, g.nacaPolygon(4 * c, H / 2 - 0.5, c, 0.12, 0));
  s.setSdf(sdf);
  s.initField();
  const forces = [];
  let prev = Infinity;
  let cd = NaN;
  for (let k = 0; k < 30; k++) {
    await runRamped(s, 10000, u, 5000, forces);
    cd = (2 * g.mean(forces.slice(-20).map((f) => f.fx))) / (u * u * c);
    if (k >= 5 && Math.abs(cd - prev) < 2e-4) break;
    prev = cd;
  }
  const steps = s.step;
  s.destroy();
  return { name: 'NACA0012 (Re 500, alpha 0)', metrics: [metric('C_D', cd, [0.1762, 0.178], around(0.1762, 0.05))],
    steps, cells: W * H, notes: [`chord ${c} cells, domain 16c x 8c`] };
}

/** src/cases3d.ts:11-81: sphere C_D at Re 100 in FP32 and FP16 storage, 192 x 96 x 96. */
async function sphereFp16() {
  const [W, H, D] = [192, 96, 96];
  const [d, u] = [16, 0.1];
  const tau = g.deriveTau(100, u, d).tau;
  const sdf = g.emptySdf3(W, H, D);
  g.addSphere(sdf, W, H, D, W / 4, H / 2, D / 2, d / 2);
  const scale = 2 / (u * u * g.referenceArea('sphere', d, D));
  const cd = {};
  const notes = [];
  let steps = 0;
  for (const precision of ['fp32', 'fp16']) {
    const s = await Solver3D.create(0, { width: W, height: H, depth: D, precision, tau, spongeFraction: 0.15,
      spongeTau: 1, absorb: 0.02, forceEvery: 4 });
    s.setSdf(sdf);
    s.initField();
    let prev = NaN;
    let converged = false;
    for (let k = 1; k < 16; k++) {
      const block = [];
      const end = k * 2000;
      while (s.step < end) {
        const r = Math.min(1, s.step / 3000);
        s.setInlet(u * r * r * (3 - 2 * r));
        await s.run(Math.min(400, end - s.step), 400);
        for (const f of await s.readForces()) block.push(f.fx);
      }
      const c = g.mean(block) * scale;
      cd[precision] = c;
      if (k >= 5 && Math.abs(c - prev) / Math.abs(prev) < 1e-4) {
        converged = true;
        break;
      }
      prev = c;
    }
    if (!converged) notes.push(`${precision} reached 15 blocks without converging`);
    steps += s.step;
    s.destroy();
  }
  notes.push(`C_D FP32 ${cd.fp32.toFixed(5)}, FP16 ${cd.fp16.toFixed(5)}`);
  return { name: 'Sphere C_D, FP16 vs FP32 (Re 100)',
    metrics: [metric('C_D relative difference FP16 vs FP32', Math.abs(cd.fp16 - cd.fp32) / cd.fp32, [0, 0], [0, 0.01])],
    steps, cells: W * H * D, notes };
}

const CASES = { poiseuille, periodicSeam, taylorGreen, schaferTurek1, schaferTurek2, unconfinedCylinder, cavity, naca,
  sphereFp16 };

async function runCase(name) {
  const t0 = Date.now();
  const r = await CASES[name]();
  r.seconds = (Date.now() - t0) / 1000;
  r.pass = r.metrics.length > 0 && r.metrics.every((m) => m.pass);
  return r;
}

module.exports = { CASES, runCase, runRamped, metric };

if (require.main === module) {
  const names = process.argv.slice(2);
  (async () => {
    let ok = true;
    for (const n of names.length ? names : Object.keys(CASES)) {
      if (!CASES[n]) throw new Error(`unknown case ${n}`);
      const r = await runCase(n);
      ok = ok && r.pass;
      console.log(JSON.stringify(r));
    }
    process.exit(ok ? 0 : 1);
  })().catch((e) => {
    console.log(`FAIL ${e.stack || e}`);
    process.exit(1);
  });
}
