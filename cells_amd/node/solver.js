'use strict';
/**
 * Solver (D2Q9) and Solver3D (D3Q19) for Node.js over the gfx950 library, with the method surface
 * of the reference's WebGPU classes (src/solver.ts:50-340, src/solver3d.ts:42-319): create,
 * update / setTau / setSmagorinsky / setInlet, setSdf, initField, resetForces, encodeSteps,
 * encodeMacro, run, readMacro, readForces, destroy, and the readonly W, H, D, N, step, cfg.
 *
 * `device` is a GPU ordinal (or anything else for GPU 0) instead of a GPUDevice. encodeSteps /
 * encodeMacro take the pass argument for source compatibility and ignore it: work is enqueued on
 * the solver's own HIP stream, in order, exactly as the reference's single queue orders it.
 * Written for Node >= 12 (no ?? or ?.).
 */
const path = require('path');

let native = null;
function addon() {
  if (!native) native = require(path.join(__dirname, 'lbm_napi.node'));
  return native;
}

const XMode = Object.freeze({ Periodic: 0, Wall: 1, Open: 2 });
const YMode = Object.freeze({ Periodic: 0, NoSlip: 1, Slip: 2, MovingWall: 3 });
const MAGIC_LAMBDA = 3 / 16;
const HISTORY_LEN = 8192;
const PRECISION_CODES = Object.freeze({ fp32: 0, fp16: 1, bf16: 2, fp16sr: 3 });
const PRESETS3 = Object.freeze({ low: [128, 64, 64], medium: [192, 96, 96], high: [256, 128, 128] });

function deviceOrdinal(device) {
  return typeof device === 'number' && device >= 0 ? Math.floor(device) : 0;
}

/** trtLambda3 (src/lattice3d.ts:29-31): 3/16 where stable, shrinking toward BGK near tau = 1/2. */
function trtLambda3(tau) {
  return Math.min(MAGIC_LAMBDA, 50 * (tau - 0.5) * (tau - 0.5));
}

/** Bytes one distribution buffer needs and whether it fits the given limits (src/shaders/common3d.ts:107-110). */
function bytesPerPopulation(p) {
  return p === 'fp32' ? 4 : 2;
}
function fitsLimits(W, H, D, p, limits) {
  const bytes = 19 * bytesPerPopulation(p) * W * H * D;
  return bytes <= limits.maxStorageBufferBindingSize && bytes <= limits.maxBufferSize;
}

function flat2(c) {
  const bf = c.bodyForce || [0, 0];
  return {
    width: c.width, height: c.height, left: c.left, right: c.right, bottom: c.bottom, top: c.top,
    parabolic: c.parabolic ? 1 : 0, workgroupSize: c.workgroupSize, forceEvery: c.forceEvery, tau: c.tau,
    lambda: c.lambda, smagorinsky: c.smagorinsky, uIn: c.uIn, uLid: c.uLid, spongeFraction: c.spongeFraction,
    spongeTau: c.spongeTau, bodyForceX: bf[0], bodyForceY: bf[1], absorb: c.absorb,
    inletLayerFraction: c.inletLayerFraction,
  };
}

function flat3(c) {
  const o = {
    width: c.width, height: c.height, depth: c.depth, precisionCode: PRECISION_CODES[c.precision],
    workgroupSize: c.workgroupSize, forceEvery: c.forceEvery, tau: c.tau, smagorinsky: c.smagorinsky,
    uIn: c.uIn, spongeFraction: c.spongeFraction, spongeTau: c.spongeTau, absorb: c.absorb,
    velocitySet: c.velocitySet === undefined ? 19 : c.velocitySet,
  };
  if (c.lambda !== undefined && c.lambda !== null) o.lambda = c.lambda;
  return o;
}

/**
 * Shared plumbing: async waits are tracked so destroy() can wait for them before freeing, and while a
 * ShapeAdjoint pass runs on the worker pool (it owns the solver's buffers and stream) every other
 * method throws instead of racing it.
 */
class Base {
  constructor() {
    this._pending = new Set();
    this._busy = null;
    this._h = null;
  }
  get handle() {
    if (this._busy) throw new Error(`${this._busy} is running on this solver`);
    return this._h;
  }
  set handle(h) {
    this._h = h;
  }
  /**
   * Runs start(rawHandle) as the solver's only user: after the waits already pending settle, and until
   * the promise start returns settles.
   */
  _exclusive(name, start) {
    const h = this.handle;
    if (!h) throw new Error('solver destroyed');
    this._busy = name;
    const before = Array.from(this._pending).map((q) => q.catch(() => undefined));
    const p = Promise.all(before).then(() => start(h));
    const clear = () => {
      this._busy = null;
    };
    p.then(clear, clear);
    return this._track(p);
  }
  _track(p) {
    this._pending.add(p);
    const done = () => this._pending.delete(p);
    p.then(done, done);
    return p;
  }
  get step() {
    return addon().info(this.handle).step;
  }
  setSdf(sdf) {
    if (!(sdf instanceof Float32Array)) sdf = Float32Array.from(sdf);
    addon().setSdf(this.handle, sdf);
  }
  resetForces() {
    addon().resetForces(this.handle);
  }
  /** Inlet ramp evaluated per step on the GPU side: 'off' | 'linear' | 'smoothstep' (an extension). */
  setInletRamp(kind, uTarget, steps) {
    const k = { off: 0, linear: 1, smoothstep: 2 }[kind];
    if (k === undefined) throw new Error(`unknown ramp ${kind}`);
    addon().setInletRamp(this.handle, k, uTarget, steps);
  }
  encodeSteps(pass, n) {
    if (n === undefined && typeof pass === 'number') n = pass;
    addon().enqueueSteps(this.handle, n);
  }
  encodeMacro() {
    addon().enqueueMacro(this.handle);
  }
  /** Runs n steps and refreshes the macro buffer; resolves when the GPU has finished (src/solver.ts:295-307). */
  run(n, chunk) {
    const c = chunk === undefined ? 500 : chunk;
    for (let done = 0; done < n; ) {
      const k = Math.min(c, n - done);
      this.encodeSteps(null, k);
      this.encodeMacro();
      done += k;
    }
    return this._track(addon().syncAsync(this.handle));
  }
  readMacro() {
    return this._track(addon().readMacroAsync(this.handle));
  }
  /** Force samples since the last call, oldest first; a read overtaken by resetForces resolves to []. */
  readForces() {
    return this._track(addon().readForcesAsync(this.handle));
  }
  /** Current distribution buffer in storage format (Float32Array, or Uint16Array of f16 / bf16 words). */
  readPopulations() {
    return addon().readPopulations(this.handle);
  }
  readFlags() {
    return addon().readFlags(this.handle);
  }
  /** Milliseconds of n steps between two HIP events on the solver's stream. */
  timeSteps(n) {
    return addon().timeSteps(this.handle, n);
  }
  /**
   * hasDiverged (src/tunnel3d.ts:97-99) kept on the device: the number of the first step whose force
   * sample had a non-finite component since the last initField / resetForces, 0 while none had.
   * Checked as the samples are written, so a blow-up between two readForces calls is not missed.
   * Waits for the work enqueued before it.
   */
  divergedAt() {
    return addon().divergedAt(this.handle);
  }
  /** Picks the fastest launch shape (and, for 3D grids, device population layout) for this grid
   * (results do not change); returns { block, sweep, tiled, stride }. */
  autotune(trialSteps) {
    return addon().autotune(this.handle, trialSteps === undefined ? 200 : trialSteps);
  }
  /** Device population layout by hand: tiles of 64 cells, or one array per population at a stride
   * (0: unpadded). Results and every array read back are unchanged. */
  setPopulationLayout(tiled, stride) {
    addon().setPopulationLayout(this.handle, !!tiled, stride === undefined ? 0 : stride);
  }
  destroy() {
    const h = this._h;
    if (!h) return;
    this._h = null;
    if (this._pending.size === 0) addon().destroy(h);
    else Promise.all(Array.from(this._pending).map((p) => p.catch(() => undefined))).then(() => addon().destroy(h));
  }
}

class Solver extends Base {
  /** Solver.create (src/solver.ts:87-92). */
  static async create(device, config) {
    return new Solver(device, config);
  }
  constructor(device, config) {
    super();
    this.cfg = Object.assign(
      {
        lambda: MAGIC_LAMBDA, smagorinsky: 0, uIn: 0, parabolic: false, uLid: 0, spongeFraction: 0, spongeTau: 1,
        bodyForce: [0, 0], absorb: 0, inletLayerFraction: 0, workgroupSize: 128, forceEvery: 1,
      },
      config,
    );
    this.W = config.width;
    this.H = config.height;
    this.N = this.W * this.H;
    this.wg = this.cfg.workgroupSize;
    this.handle = addon().create2d(flat2(this.cfg), deviceOrdinal(device));
  }
  /** update (src/solver.ts:175-178). */
  update(changes) {
    Object.assign(this.cfg, changes);
    addon().update2d(this.handle, flat2(this.cfg));
  }
  /** initField (src/solver.ts:220-234): equilibrium at (rho, ux, uy) per cell, or at rest. */
  initField(macro) {
    addon().initField(this.handle, macro || null);
  }
  /** rebuildFlags (src/solver.ts:205-217). */
  rebuildFlags() {
    addon().rebuildFlags(this.handle);
  }
  /** brush (src/solver.ts:249-268): union (or, erasing, subtract) a disc into the SDF. */
  brush(cx, cy, r, erase, rebuild) {
    addon().brush(this.handle, cx, cy, r, !!erase, rebuild === undefined ? true : !!rebuild);
  }
  /** brushStroke (src/solver.ts:243-247): a disc at every [x, y] point, then one rebuild. */
  brushStroke(points, r, erase) {
    const xy = new Float64Array(points.length * 2);
    points.forEach((p, k) => {
      xy[2 * k] = p[0];
      xy[2 * k + 1] = p[1];
    });
    addon().brushStroke(this.handle, xy, r, !!erase);
  }
}

class Solver3D extends Base {
  /** Solver3D.create (src/solver3d.ts:79-104): allocation failures reject with 'Solver3D allocation failed: ...'. */
  static async create(device, config) {
    if (PRECISION_CODES[config.precision] === undefined) throw new Error(`unknown precision ${config.precision}`);
    return new Solver3D(device, config);
  }
  constructor(device, config) {
    super();
    this.cfg = Object.assign(
      { smagorinsky: 0, uIn: 0, spongeFraction: 0, spongeTau: 1, absorb: 0, workgroupSize: 128, forceEvery: 1 },
      config,
    );
    this.W = config.width;
    this.H = config.height;
    this.D = config.depth;
    this.N = this.W * this.H * this.D;
    this.wg = this.cfg.workgroupSize;
    this.handle = addon().create3d(flat3(this.cfg), deviceOrdinal(device));
  }
  _push() {
    addon().update3d(this.handle, flat3(this.cfg));
  }
  setTau(tau) {
    this.cfg.tau = tau;
    this._push();
  }
  setSmagorinsky(cs) {
    this.cfg.smagorinsky = cs;
    this._push();
  }
  setInlet(u) {
    this.cfg.uIn = u;
    this._push();
  }
  /**
   * Open x planes (an extension): 'equilibrium', the reference's (src/shaders/step3d.ts:190-209), or
   * 'zouHe', Hecht-Harting velocity inlet and rho = 1 pressure outlet whose nodes then collide.
   */
  setOpenFaces(kind) {
    const k = { equilibrium: 0, zouHe: 1 }[kind];
    if (k === undefined) throw new Error(`unknown open faces ${kind}`);
    addon().setOpenFaces(this.handle, k);
  }
  /** initField (src/solver3d.ts:231-242): inflow equilibrium in fluid cells, rest in solids. */
  initField() {
    addon().initField(this.handle, null);
  }
}

const RAMP_CODES = Object.freeze({ off: 0, none: 0, linear: 1, smoothstep: 2 });

/**
 * Shape gradient of a drag objective through the discrete adjoint (no counterpart in the reference):
 * forward(sdf) runs the tunnel from its initial field for `steps` steps and resolves to
 * J = scale * mean over the last `window` force samples of direction . F; backward(seed) resolves to
 * seed * dJ/dsdf at every node (Float64Array of N). Both run on the libuv pool over
 * lbm_adjoint_forward / lbm_adjoint_backward.
 */
class ShapeAdjoint {
  constructor(solver, opts) {
    const o = opts || {};
    const ramp = o.ramp || null;
    const kind = ramp ? RAMP_CODES[ramp.kind] : 0;
    if (kind === undefined) throw new Error(`unknown ramp ${ramp.kind}`);
    if (!(o.steps > 0) || !(o.window > 0) || o.window > o.steps) throw new Error('need 0 < window <= steps');
    this.solver = solver;
    this.opts = {
      steps: o.steps, windowStart: o.steps - o.window, direction: (o.direction || [1, 0, 0]).concat([0, 0, 0]).slice(0, 3),
      scale: o.scale === undefined ? 1 : o.scale, checkpointEvery: o.checkpointEvery || 200, rampKind: kind,
      rampU: ramp ? ramp.uTarget : 0, rampSteps: ramp ? ramp.steps : 0, recordBudgetBytes: o.recordBudgetBytes || 0,
    };
  }
  forward(sdf) {
    if (sdf !== undefined && sdf !== null) this.solver.setSdf(sdf);
    return this.solver._exclusive('ShapeAdjoint.forward', (h) => addon().adjointForwardAsync(h, this.opts));
  }
  backward(seed) {
    const k = seed === undefined ? 1 : seed;
    return this.solver._exclusive('ShapeAdjoint.backward', (h) => addon().adjointBackwardAsync(h, k));
  }
}

module.exports = {
  Solver, Solver3D, ShapeAdjoint, XMode, YMode, MAGIC_LAMBDA, HISTORY_LEN, PRESETS3, trtLambda3, fitsLimits,
  bytesPerPopulation, addon,
};
