'use strict';
// Node wrapper on the GPU: Solver / Solver3D through the N-API addon and the C ABI, including the
// readForces-across-resetForces semantics the reference pins (src/solver.test.ts:38-81,
// src/solver3d.test.ts:157-195) and its allocation-failure message (src/solver3d.ts:83-86).
const assert = require('assert');
const path = require('path');
const m = require(path.join(__dirname, '..', '..', 'cells_amd', 'node'));

async function main() {
  // --- 3D tunnel: run, forces, macro
  const [W, H, D] = [32, 16, 14];
  const s3 = await m.Solver3D.create(0, { width: W, height: H, depth: D, precision: 'fp32', tau: 0.524, uIn: 0.05,
    spongeFraction: 0.15, spongeTau: 1, absorb: 0.02, forceEvery: 1 });
  const sdf = m.emptySdf3(W, H, D);
  m.addSphere(sdf, W, H, D, 8, 7.5, 6.5, 3);
  s3.setSdf(sdf);
  s3.initField();
  await s3.run(50, 20);
  assert.strictEqual(s3.step, 50);
  const f = await s3.readForces();
  assert.strictEqual(f.length, 50);
  assert.strictEqual(f[0].step, 1);
  assert.ok(f[49].fx > 0 && Number.isFinite(f[49].fz));
  const mac = await s3.readMacro();
  assert.strictEqual(mac.length, W * H * D * 4);
  assert.ok(!m.fieldDiverged(mac));
  assert.strictEqual(s3.divergedAt(), 0);
  console.log(`ok 3d run: C_D-like fx ${f[49].fx.toExponential(4)}`);

  // --- blow-up record (hasDiverged kept on the device): a 2D tunnel driven unstable records the step
  // of its first non-finite force sample, the same step hasDiverged finds in the samples
  {
    const u = await m.Solver.create(0, { width: 64, height: 24, left: m.XMode.Open, right: m.XMode.Open,
      bottom: m.YMode.Slip, top: m.YMode.Slip, tau: 0.5005, uIn: 0.3, spongeFraction: 0, absorb: 0, forceEvery: 1 });
    const sdf2 = m.emptySdf(64, 24);
    m.addCircle(sdf2, 64, 24, 16, 11.5, 4);
    u.setSdf(sdf2);
    u.initField();
    await u.run(300, 100);
    const samples = await u.readForces();
    assert.ok(m.hasDiverged(samples));
    const firstBad = samples.find((q) => !Number.isFinite(q.fx) || !Number.isFinite(q.fy)).step;
    assert.strictEqual(u.divergedAt(), firstBad);
    u.resetForces();
    assert.strictEqual(u.divergedAt(), 0);
    u.destroy();
    console.log(`ok blow-up record at step ${firstBad}`);
  }

  // --- Zou-He faces: the inlet plane carries u = (uIn, 0, 0) (no absorbing source at x = 0)
  s3.setOpenFaces('zouHe');
  await s3.run(50, 25);
  const mz = await s3.readMacro();
  let du = 0;
  let drho = 0;
  for (let k = 0; k < H * D; k++) {
    const c = k * W * 4;
    du = Math.max(du, Math.abs(mz[c] - 0.05), Math.abs(mz[c + 1]), Math.abs(mz[c + 2]));
    drho = Math.max(drho, Math.abs(mz[c + 3] - 1));
  }
  assert.ok(du < 1e-6 && drho > 1e-6, `${du} ${drho}`);
  assert.throws(() => s3.setOpenFaces('open'), /unknown open faces/);
  s3.setOpenFaces('equilibrium');

  // --- D3Q27 velocity set through the same wrapper
  const q27 = await m.Solver3D.create(0, { width: W, height: H, depth: D, precision: 'fp32', tau: 0.524, uIn: 0.05,
    spongeFraction: 0.15, spongeTau: 1, absorb: 0.02, forceEvery: 1, velocitySet: 27 });
  q27.setSdf(sdf);
  q27.initField();
  await q27.run(50, 25);
  const f27 = await q27.readForces();
  assert.strictEqual(f27.length, 50);
  assert.ok(Math.abs(f27[49].fx / f[49].fx - 1) < 0.2, `${f27[49].fx} ${f[49].fx}`);
  assert.strictEqual((await q27.readPopulations()).length, 27 * W * H * D);
  assert.throws(() => q27.setOpenFaces('zouHe'), /D3Q19 only/);
  q27.destroy();
  console.log(`ok d3q27: fx ${f27[49].fx.toExponential(4)} against d3q19 ${f[49].fx.toExponential(4)}`);
  console.log(`ok zou-he faces: inlet |u - (uIn, 0, 0)| ${du.toExponential(2)}, |rho - 1| ${drho.toExponential(2)}`);

  // --- stale read across resetForces (2D and 3D)
  for (const dim of [2, 3]) {
    const s = dim === 2
      ? await m.Solver.create(0, { width: 48, height: 20, left: m.XMode.Open, right: m.XMode.Open, bottom: m.YMode.Slip,
        top: m.YMode.Slip, tau: 0.6, uIn: 0.04, forceEvery: 4 })
      : await m.Solver3D.create(0, { width: 24, height: 12, depth: 10, precision: 'fp16', tau: 0.56, uIn: 0.04, forceEvery: 4 });
    const g = dim === 2 ? m.emptySdf(48, 20) : m.emptySdf3(24, 12, 10);
    if (dim === 2) m.addCircle(g, 48, 20, 12, 9.5, 3); else m.addSphere(g, 24, 12, 10, 7, 5.5, 4.5, 2);
    s.setSdf(g);
    s.initField();
    await s.run(4000, 1000);
    const stale = s.readForces();
    s.initField(); // step = 0 and resetForces
    const r1 = s.run(40);
    assert.deepStrictEqual(await stale, []);
    await r1;
    await s.run(400);
    const samples = await s.readForces();
    assert.strictEqual(samples.length, 110);
    assert.strictEqual(samples[0].step, 4);
    assert.strictEqual(samples[109].step, 440);
    await s.run(400);
    assert.strictEqual((await s.readForces())[0].step, 444);
    s.destroy();
    s.destroy(); // idempotent
    console.log(`ok ${dim}d force ring semantics`);
  }

  // --- shape gradient through ShapeAdjoint: dJ/dcx and dJ/dr of a circle against central differences
  {
    const [AW, AH, u, cx, cy, r] = [64, 24, 0.05, 16.3, 11.5, 4];
    const sa = await m.Solver.create(0, { width: AW, height: AH, left: m.XMode.Open, right: m.XMode.Open,
      bottom: m.YMode.Slip, top: m.YMode.Slip, tau: 0.56, uIn: u, spongeFraction: 0.15, spongeTau: 1, absorb: 0.02,
      forceEvery: 1 });
    const adj = new m.ShapeAdjoint(sa, { steps: 200, window: 80, direction: [1, 0], scale: 2 / (u * u * 2 * r),
      checkpointEvery: 50, ramp: { kind: 'smoothstep', uTarget: u, steps: 60 } });
    const circle = (x, rad) => {
      const g = m.emptySdf(AW, AH);
      m.addCircle(g, AW, AH, x, cy, rad);
      return g;
    };
    const sdf0 = circle(cx, r);
    const J0 = await adj.forward(sdf0);
    const grad = await adj.backward();
    assert.strictEqual(grad.length, AW * AH);
    assert.strictEqual(await adj.forward(circle(cx, r)), J0);
    const g2 = await adj.backward(-2);
    // dsdf/dcx = -(x - cx) / |x - c| and dsdf/dr = -1 where the circle wrote the SDF, 0 elsewhere
    let dcx = 0;
    let dr = 0;
    const gmax = grad.reduce((a, v) => Math.max(a, Math.abs(v)), 0);
    assert.ok(gmax > 0);
    for (let y = 0; y < AH; y++) {
      for (let x = 0; x < AW; x++) {
        const k = y * AW + x;
        assert.ok(Math.abs(g2[k] + 2 * grad[k]) <= 2e-5 * Math.abs(grad[k]) + 2e-6 * gmax, 'seed scaling');
        if (sdf0[k] >= m.FAR) continue;
        const dist = Math.hypot(x - cx, y - cy);
        if (dist > 0) dcx -= grad[k] * (x - cx) / dist;
        dr -= grad[k];
      }
    }
    const h = 1e-3;
    const fdx = ((await adj.forward(circle(cx + h, r))) - (await adj.forward(circle(cx - h, r)))) / (2 * h);
    const fdr = ((await adj.forward(circle(cx, r + h))) - (await adj.forward(circle(cx, r - h)))) / (2 * h);
    console.log(`adjoint J ${J0.toFixed(6)}: dJ/dcx ${dcx.toExponential(4)} (FD ${fdx.toExponential(4)}), ` +
      `dJ/dr ${dr.toExponential(4)} (FD ${fdr.toExponential(4)})`);
    const scale = Math.max(Math.abs(fdx), Math.abs(fdr));
    assert.ok(Math.abs(dcx - fdx) <= 0.02 * Math.abs(fdx) + 2e-3 * scale, 'dJ/dcx');
    assert.ok(Math.abs(dr - fdr) <= 0.02 * Math.abs(fdr) + 2e-3 * scale, 'dJ/dr');
    sa.destroy();
    console.log('ok shape adjoint');
  }

  // --- device population layout: tiles of 64 cells give the same bits; autotune reports the layout
  {
    const mk = async () => {
      const t = await m.Solver3D.create(0, { width: 64, height: 16, depth: 12, precision: 'fp32', tau: 0.524, uIn: 0.05,
        spongeFraction: 0.15, spongeTau: 1, absorb: 0.02, forceEvery: 1 });
      const sd = m.emptySdf3(64, 16, 12);
      m.addSphere(sd, 64, 16, 12, 16, 7.5, 5.5, 3);
      t.setSdf(sd);
      t.initField();
      return t;
    };
    const a = await mk();
    const b = await mk();
    b.setPopulationLayout(true);
    await a.run(30, 10);
    await b.run(30, 10);
    assert.deepStrictEqual(await b.readPopulations(), await a.readPopulations());
    assert.throws(() => b.setPopulationLayout(false, 7), /stride/);
    const pick = a.autotune(10);
    assert.ok(typeof pick.tiled === 'boolean' && pick.stride > 0, JSON.stringify(pick));
    a.destroy();
    b.destroy();
    console.log('ok population layout');
  }

  // --- allocation failure message keeps the reference's prefix
  let err = null;
  try {
    await m.Solver3D.create(0, { width: 4096, height: 4096, depth: 16, precision: 'fp32', tau: 0.6 });
  } catch (e) {
    err = e;
  }
  assert.ok(err && /^Solver3D allocation failed: /.test(err.message), err && err.message);
  console.log('ok allocation failure message');
  s3.destroy();
  console.log('all passed');
}

main().catch((e) => {
  console.log(`FAIL ${e.stack || e}`);
  process.exit(1);
});
