'use strict';
// Host-side checks of the Node wrapper that need no GPU: the reference's own unit-test values
// (src/analysis.test.ts, src/geometry3d.test.ts, src/shaders/common3d.test.ts, src/lattice3d.test.ts).
const assert = require('assert');
const path = require('path');
const m = require(path.join(__dirname, '..', '..', 'cells_amd', 'node'));

let failures = 0;
function test(name, fn) {
  try {
    fn();
    console.log(`ok ${name}`);
  } catch (e) {
    failures++;
    console.log(`FAIL ${name}: ${e.message}`);
  }
}
const close = (a, b, tol, msg) => assert.ok(Math.abs(a - b) <= tol, `${msg || ''} ${a} vs ${b}`);

test('periodFromCrossings recovers a sampled sine', () => {
  const t = [], v = [];
  for (let k = 0; k < 5000; k++) { t.push(k * 0.7); v.push(0.3 + Math.sin((2 * Math.PI * k * 0.7) / 123.4)); }
  close(m.periodFromCrossings(t, v).period, 123.4, 0.005);
  assert.strictEqual(m.periodFromCrossings([0, 1, 2], [1, 2, 3]), null);
});

test('deriveTau and clamp', () => {
  close(m.deriveTau(100, 0.1, 40).tau, 0.62, 1e-10);
  const c = m.deriveTau(1e5, 0.1, 40);
  assert.ok(c.clamped);
  assert.strictEqual(c.tau, 0.51);
  close(c.effectiveRe, m.maxReynolds(0.1, 40), 1e-6);
});

test('circle and naca sdf', () => {
  const s = m.emptySdf(20, 20);
  m.addCircle(s, 20, 20, 10, 10, 4);
  close(s[10 * 20 + 10], -4, 1e-6);
  close(s[10 * 20 + 16], 2, 1e-6);
  const W = 200, H = 60, sd = m.emptySdf(W, H);
  const poly = m.nacaPolygon(50, 30, 100, 0.12, 0);
  m.addPolygon(sd, W, H, poly);
  assert.ok(sd[30 * W + 80] < -5);
  assert.ok(sd[30 * W + 40] > 9);
  const ys = poly.map((p) => p[1]);
  close(Math.max.apply(null, ys) - Math.min.apply(null, ys), 12, 0.05);
});

test('sphere, cylinder, referenceArea', () => {
  const [W, H, D] = [32, 24, 20];
  const at = (sdf, x, y, z) => sdf[x + W * (y + H * z)];
  const s = m.emptySdf3(W, H, D);
  m.addSphere(s, W, H, D, 16, 12, 10, 4);
  close(at(s, 16, 12, 10), -4, 1e-6);
  close(at(s, 22, 12, 10), 2, 1e-6);
  assert.strictEqual(at(s, 29, 12, 10), m.FAR);
  const c = m.emptySdf3(W, H, D);
  m.addCylinderZ(c, W, H, D, 16, 12, 3);
  close(at(c, 16, 12, 0), -3, 1e-6);
  close(at(c, 16, 12, D - 1), -3, 1e-6);
  close(m.referenceArea('sphere', 16, 96), 201.06, 0.005);
  assert.strictEqual(m.referenceArea('cube', 16, 96), 256);
  assert.strictEqual(m.referenceArea('cylinder', 16, 96), 1536);
  close(m.referenceArea('wing', 16, 96), 921.6, 1e-6);
  assert.strictEqual(m.referenceArea('none', 16, 96), 0);
});

test('wing tips are flat', () => {
  const [W, H, D] = [32, 24, 20];
  const at = (sdf, x, y, z) => sdf[x + W * (y + H * z)];
  const s = m.emptySdf3(W, H, D);
  const [cx, cy, cz, chord, span] = [16, 12, 10, 16, 12];
  m.addWing(s, W, H, D, cx, cy, cz, chord, span, 0);
  const x = Math.round(cx - chord / 2 + 0.3 * chord);
  close(at(s, x, cy, cz + span / 2 + 1), 1, 0.05);
  close(at(s, x, cy, cz - span / 2 - 1), 1, 0.05);
  assert.ok(at(s, x, cy, cz) < 0);
});

test('fitsLimits truth table', () => {
  const small = { maxStorageBufferBindingSize: 128 * 2 ** 20, maxBufferSize: 256 * 2 ** 20 };
  const fits = (k, p) => m.fitsLimits(m.PRESETS3[k][0], m.PRESETS3[k][1], m.PRESETS3[k][2], p, small);
  assert.deepStrictEqual([fits('low', 'fp16'), fits('low', 'fp32'), fits('medium', 'fp16'), fits('medium', 'fp32'),
    fits('high', 'fp16'), fits('high', 'fp32')], [true, true, true, false, false, false]);
});

test('trtLambda3 known answers', () => {
  assert.strictEqual(m.trtLambda3(0.6), 3 / 16);
  close(m.trtLambda3(0.5288), 50 * 0.0288 ** 2, 1e-12);
  close(m.trtLambda3(0.51), 0.005, 1e-12);
});

test('divergence checks', () => {
  assert.ok(m.hasDiverged([{ fx: 1, fy: NaN }]));
  assert.ok(!m.hasDiverged([{ fx: 1, fy: 2, fz: 3 }]));
  assert.ok(m.fieldDiverged(new Float32Array([0, Infinity])));
  close(m.coefficientScale(0.1, 201.06), 2 / (0.01 * 201.06), 1e-12);
});

test('addon exports every entry point', () => {
  const a = m.addon();
  for (const k of ['create2d', 'create3d', 'destroy', 'update2d', 'update3d', 'setSdf', 'initField', 'resetForces',
    'enqueueSteps', 'enqueueMacro', 'syncAsync', 'readMacroAsync', 'readForcesAsync', 'readPopulations', 'readFlags',
    'info', 'timeSteps', 'setInletRamp', 'adjointForwardAsync', 'adjointBackwardAsync', 'rebuildFlags', 'setOpenFaces', 'brush', 'brushStroke', 'deviceName', 'copyRoofline', 'autotune', 'divergedAt']) assert.strictEqual(typeof a[k], 'function', k);
  assert.strictEqual(a.abiVersion(), 2);
});

// An exclusive pass (ShapeAdjoint) starts after the pending waits and locks the solver until it settles.
async function exclusiveGuard() {
  const s = Object.create(m.Solver.prototype);
  s._pending = new Set();
  s._busy = null;
  s._h = { fake: true };
  const order = [];
  s._track(new Promise((r) => setTimeout(() => { order.push('read'); r(); }, 20)));
  const p = s._exclusive('ShapeAdjoint.forward', (h) => {
    order.push('pass');
    assert.strictEqual(h, s._h);
    return Promise.resolve(1.5);
  });
  assert.throws(() => s.handle, /ShapeAdjoint.forward is running on this solver/);
  assert.strictEqual(await p, 1.5);
  assert.deepStrictEqual(order, ['read', 'pass']);
  assert.strictEqual(s.handle, s._h);
  const q = s._exclusive('ShapeAdjoint.backward', () => Promise.reject(new Error('boom')));
  await assert.rejects(q, /boom/);
  assert.strictEqual(s.handle, s._h);
}

exclusiveGuard().then(
  () => console.log('ok exclusive adjoint pass'),
  (e) => {
    failures++;
    console.log(`FAIL exclusive adjoint pass: ${e.message}`);
  },
).then(() => {
  if (failures) {
    console.log(`${failures} failed`);
    process.exit(1);
  }
  console.log('all passed');
});
