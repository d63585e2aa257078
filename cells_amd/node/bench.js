'use strict';
/**
 * The reference's benchmark page (src/bench.ts:61-186) on the gfx950 library: for each 2D grid a copy
 * roofline and the periodic and tunnel+cylinder scenarios at workgroups 64/128/256, then the D3Q19
 * medium grid, empty and with a sphere, per storage precision, median of 3, and the spec floors.
 * Rows carry the same fields and log lines as BenchRow and its log format; each line adds the
 * fraction of MI355X's 8 TB/s HBM peak that the row's algorithmic GB/s reaches. Timing is HIP events
 * on the solver's stream around the enqueued steps (the role of the reference's timestamp queries).
 *
 * Additions, marked in the scenario name or workgroup column: D3Q19 rows in bf16 and, per 3D
 * scenario, a row at the autotuned launch shape ("auto").
 *
 *   node cells_amd/node/bench.js [--quick] [--json out.json]
 */
const { Solver, Solver3D, XMode, YMode, PRESETS3, bytesPerPopulation, addon } = require('./solver');
const { emptySdf, addCircle, emptySdf3, addSphere } = require('./geometry');

/** Bytes moved per cell update: nine loads and nine stores of f32 plus the flag word (src/bench.ts:24). */
const BYTES_PER_CELL = 9 * 4 * 2 + 4;
/** D3Q19: nineteen loads and nineteen stores of b bytes plus the flag word (src/bench.ts:27). */
const bytesPerCell3D = (p) => 19 * bytesPerPopulation(p) * 2 + 4;
/** Spec throughput floors of the reference, for the empty medium grid at workgroup 128 (src/bench.ts:30). */
const FLOOR_3D = { fp16: 1200, fp32: 600 };
const HBM_PEAK_GBS = 8000;
const TIMER = 'gpu-timestamps';

const peak = (gbps) => `  ${((100 * gbps) / HBM_PEAK_GBS).toFixed(1).padStart(5)}% of HBM peak`;

/** Warm-up of 50 steps, then the timed steps (src/bench.ts:34-58, 128-129). */
function measureSteps(s, steps) {
  s.timeSteps(50);
  return s.timeSteps(steps);
}

function copyRoofline(W, H, wg, steps) {
  return addon().copyRoofline(0, W * H, 9, 4, wg, steps);
}

async function runBenchmark(log, opts) {
  const o = opts || {};
  const say = log || (() => {});
  const adapter = addon().deviceName(0);
  say(`adapter: ${adapter}; timer: HIP events`);
  const rows = [];
  const sizes = o.quick ? [[512, 256]] : [[512, 256], [1024, 512], [2048, 1024], [4096, 2048]];
  const wgs = o.quick ? [128] : [64, 128, 256];
  const scenarios = [
    { name: 'periodic', cfg: { left: XMode.Periodic, right: XMode.Periodic, bottom: YMode.Periodic, top: YMode.Periodic },
      body: false, forceEvery: [0] },
    { name: 'tunnel+cylinder', cfg: { left: XMode.Open, right: XMode.Open, bottom: YMode.Slip, top: YMode.Slip, uIn: 0.1,
      spongeFraction: 0.15 }, body: true, forceEvery: [0, 1] },
  ];
  for (const [W, H] of sizes) {
    const steps = Math.max(100, Math.round((o.quick ? 1e8 : 4e8) / (W * H)));
    for (const wg of wgs) {
      const ms = copyRoofline(W, H, wg, steps);
      const mlups = (W * H * steps) / (ms * 1e3);
      const row = { scenario: 'copy roofline', width: W, height: H, workgroup: wg, forceEvery: 0, steps, ms, mlups,
        gbps: (mlups * 1e6 * BYTES_PER_CELL) / 1e9, timer: TIMER };
      rows.push(row);
      say(`${row.scenario.padEnd(16)} ${`${W}x${H}`.padEnd(10)} wg ${String(wg).padEnd(4)} ${''.padEnd(16)} ` +
        `${mlups.toFixed(0).padStart(6)} MLUPS  ${row.gbps.toFixed(0).padStart(4)} GB/s${peak(row.gbps)}`);
    }
    for (const sc of scenarios) {
      for (const wg of wgs) {
        for (const fe of sc.forceEvery) {
          const s = await Solver.create(0, Object.assign({ width: W, height: H, tau: 0.56, workgroupSize: wg, forceEvery: fe },
            sc.cfg));
          if (sc.body) {
            const sdf = emptySdf(W, H);
            addCircle(sdf, W, H, W / 4, H / 2, H / 10);
            s.setSdf(sdf);
          }
          s.initField();
          const ms = measureSteps(s, steps);
          const mlups = (W * H * steps) / (ms * 1e3);
          const row = { scenario: sc.name, width: W, height: H, workgroup: wg, forceEvery: fe, steps, ms, mlups,
            gbps: (mlups * 1e6 * BYTES_PER_CELL) / 1e9, timer: TIMER };
          rows.push(row);
          say(`${sc.name.padEnd(16)} ${`${W}x${H}`.padEnd(10)} wg ${String(wg).padEnd(4)} force ` +
            `${fe ? 'every step' : 'off       '} ${mlups.toFixed(0).padStart(6)} MLUPS  ${row.gbps.toFixed(0).padStart(4)} ` +
            `GB/s  (${steps} steps, ${ms.toFixed(1)} ms)${peak(row.gbps)}`);
          s.destroy();
        }
      }
    }
  }
  rows.push(...(await bench3D(say, o)));
  return { adapter, rows };
}

async function bench3D(say, o) {
  const [W, H, D] = PRESETS3.medium;
  const N = W * H * D;
  const steps = Math.max(100, Math.round((o.quick ? 5e8 : 2e9) / N));
  const rows = [];
  const precisions = o.quick ? ['fp32', 'fp16'] : ['fp32', 'fp16', 'bf16'];
  const wgs = o.quick ? [128, 'auto'] : [64, 128, 256, 'auto'];
  for (const precision of precisions) {
    for (const body of [false, true]) {
      for (const wg of wgs) {
        const forceEvery = body ? 4 : 0;
        const s = await Solver3D.create(0, { width: W, height: H, depth: D, precision, tau: 0.56, uIn: 0.1,
          spongeFraction: 0.15, absorb: 0.02, workgroupSize: wg === 'auto' ? 128 : wg, forceEvery });
        if (body) {
          const sdf = emptySdf3(W, H, D);
          addSphere(sdf, W, H, D, W / 4, H / 2, D / 2, 8);
          s.setSdf(sdf);
        }
        s.initField();
        let shape = '';
        if (wg === 'auto') {
          const t = s.autotune(200);
          shape = `, block ${t.block} rows ${['up', 'alt-hi', 'alt-lo'][t.sweep & 3]}${t.sweep & 4 ? ' merged stores' : ''}`;
        }
        s.timeSteps(50);
        const runs = [];
        for (let k = 0; k < 3; k++) runs.push(s.timeSteps(steps));
        const ms = runs.sort((a, b) => a - b)[1];
        const mlups = (N * steps) / (ms * 1e3);
        const scenario = `d3q19 ${body ? 'sphere' : 'empty'} ${precision}`;
        const row = { scenario, width: W, height: H, depth: D, workgroup: wg, forceEvery, steps, ms, mlups,
          gbps: (mlups * 1e6 * bytesPerCell3D(precision)) / 1e9, timer: TIMER };
        rows.push(row);
        say(`${scenario.padEnd(18)} ${`${W}x${H}x${D}`.padEnd(12)} wg ${String(wg).padEnd(4)} force ` +
          `${forceEvery ? 'every 4   ' : 'off       '} ${mlups.toFixed(0).padStart(6)} MLUPS  ` +
          `${row.gbps.toFixed(0).padStart(4)} GB/s  (${steps} steps, median of 3: ${ms.toFixed(1)} ms)${peak(row.gbps)}${shape}`);
        s.destroy();
      }
    }
  }
  for (const precision of Object.keys(FLOOR_3D)) {
    const r = rows.find((x) => x.scenario === `d3q19 empty ${precision}` && x.workgroup === 128);
    if (!r) continue;
    say(`d3q19 floor ${precision} (empty, wg 128): ${r.mlups.toFixed(0)} vs ${FLOOR_3D[precision]} MLUPS: ` +
      `${r.mlups >= FLOOR_3D[precision] ? 'met' : 'missed'}`);
  }
  return rows;
}

module.exports = { runBenchmark, BYTES_PER_CELL, bytesPerCell3D, FLOOR_3D };

if (require.main === module) {
  const args = process.argv.slice(2);
  const quick = args.includes('--quick');
  const j = args.indexOf('--json');
  runBenchmark((line) => console.log(line), { quick }).then(
    (res) => {
      if (j >= 0 && args[j + 1]) require('fs').writeFileSync(args[j + 1], JSON.stringify(res, null, 1));
    },
    (e) => {
      console.log(`FAIL ${e.stack || e}`);
      process.exit(1);
    },
  );
}
