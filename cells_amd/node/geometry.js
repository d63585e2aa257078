'use strict';
/**
 * Host-side inputs and post-processing for the solvers: signed distance fields (negative inside
 * solids; node (x, y[, z]) at integer coordinates, index y*W + x or x + W*(y + H*z)), lattice
 * units and the force analyses. Same shapes, margins and float32 storage as the reference's
 * src/geometry.ts, src/geometry3d.ts, src/units.ts, src/analysis.ts and src/tunnel3d.ts:75-105.
 */

const FAR = 1e6;
const MARGIN = 3;
const TAU_MIN = 0.51;

function emptySdf(W, H) {
  return new Float32Array(W * H).fill(FAR);
}

/** Min-union of d(x, y) over the integer box [x0, x1] x [y0, y1], clipped to the grid. */
function unionBox(sdf, W, H, x0f, y0f, x1f, y1f, d) {
  const x0 = Math.max(0, Math.floor(x0f));
  const y0 = Math.max(0, Math.floor(y0f));
  const x1 = Math.min(W - 1, Math.ceil(x1f));
  const y1 = Math.min(H - 1, Math.ceil(y1f));
  for (let y = y0; y <= y1; y++) {
    for (let x = x0; x <= x1; x++) {
      const k = y * W + x;
      const v = d(x, y);
      if (v < sdf[k]) sdf[k] = v;
    }
  }
}

function addCircle(sdf, W, H, cx, cy, r) {
  const m = r + MARGIN;
  unionBox(sdf, W, H, cx - m, cy - m, cx + m, cy + m, (x, y) => Math.hypot(x - cx, y - cy) - r);
}

/** Rectangle of half-extents (hx, hy), rotated counter-clockwise by angle (radians). */
function addBox(sdf, W, H, cx, cy, hx, hy, angle) {
  const a = angle || 0;
  const c = Math.cos(a);
  const s = Math.sin(a);
  const R = Math.hypot(hx, hy) + MARGIN;
  unionBox(sdf, W, H, cx - R, cy - R, cx + R, cy + R, (x, y) => {
    const qx = Math.abs((x - cx) * c + (y - cy) * s) - hx;
    const qy = Math.abs(-(x - cx) * s + (y - cy) * c) - hy;
    return Math.hypot(Math.max(qx, 0), Math.max(qy, 0)) + Math.min(Math.max(qx, qy), 0);
  });
}

/** Symmetric NACA 4-digit section (closed trailing edge, cosine spacing) with its leading edge at
 * (x0, y0), pitched nose-up by alpha radians; n points per surface. */
function nacaPolygon(x0, y0, chord, thickness, alpha, n) {
  const m = n === undefined ? 400 : n;
  const half = (t) => 5 * thickness * (0.2969 * Math.sqrt(t) - 0.126 * t - 0.3516 * t * t + 0.2843 * t * t * t - 0.1036 * t * t * t * t);
  const station = (k) => (1 - Math.cos((Math.PI * k) / m)) / 2;
  const unit = [];
  for (let k = 0; k <= m; k++) unit.push([station(k), half(station(k))]);
  for (let k = m - 1; k > 0; k--) unit.push([station(k), -half(station(k))]);
  const c = Math.cos(-alpha);
  const s = Math.sin(-alpha);
  return unit.map((p) => [x0 + chord * (p[0] * c - p[1] * s), y0 + chord * (p[0] * s + p[1] * c)]);
}

/** Signed distance to a closed polygon: nearest edge, negative inside by the even-odd rule. */
function polygonDistance(poly, x, y) {
  let best = Infinity;
  let inside = false;
  const n = poly.length;
  for (let a = 0; a < n; a++) {
    const b = a === 0 ? n - 1 : a - 1;
    const ax = poly[a][0], ay = poly[a][1];
    const ex = poly[b][0] - ax, ey = poly[b][1] - ay;
    const t = Math.max(0, Math.min(1, ((x - ax) * ex + (y - ay) * ey) / (ex * ex + ey * ey)));
    const dx = x - ax - t * ex, dy = y - ay - t * ey;
    best = Math.min(best, dx * dx + dy * dy);
    if ((ay > y) !== (poly[b][1] > y) && x < (ex * (y - ay)) / ey + ax) inside = !inside;
  }
  const d = Math.sqrt(best);
  return inside ? -d : d;
}

function bounds(poly) {
  let minX = Infinity, minY = Infinity, maxX = -Infinity, maxY = -Infinity;
  for (const p of poly) {
    minX = Math.min(minX, p[0]);
    maxX = Math.max(maxX, p[0]);
    minY = Math.min(minY, p[1]);
    maxY = Math.max(maxY, p[1]);
  }
  return [minX, minY, maxX, maxY];
}

function addPolygon(sdf, W, H, poly) {
  const b = bounds(poly);
  unionBox(sdf, W, H, b[0] - MARGIN, b[1] - MARGIN, b[2] + MARGIN, b[3] + MARGIN, (x, y) => polygonDistance(poly, x, y));
}

// ---------------------------------------------------------------- 3D

function emptySdf3(W, H, D) {
  return new Float32Array(W * H * D).fill(FAR);
}

/** Min-union over a box grown by MARGIN on every side, clipped to the grid. */
function unionBox3(sdf, W, H, D, box, d) {
  const lo = [0, 1, 2].map((k) => Math.max(0, Math.floor(box[k] - MARGIN)));
  const hi = [W, H, D].map((n, k) => Math.min(n - 1, Math.ceil(box[k + 3] + MARGIN)));
  for (let z = lo[2]; z <= hi[2]; z++) {
    for (let y = lo[1]; y <= hi[1]; y++) {
      for (let x = lo[0]; x <= hi[0]; x++) {
        const k = x + W * (y + H * z);
        const v = d(x, y, z);
        if (v < sdf[k]) sdf[k] = v;
      }
    }
  }
}

function addSphere(sdf, W, H, D, cx, cy, cz, r) {
  unionBox3(sdf, W, H, D, [cx - r, cy - r, cz - r, cx + r, cy + r, cz + r], (x, y, z) => Math.hypot(x - cx, y - cy, z - cz) - r);
}

function addBox3(sdf, W, H, D, cx, cy, cz, hx, hy, hz, angleZ) {
  const a = angleZ || 0;
  const c = Math.cos(a);
  const s = Math.sin(a);
  const R = Math.hypot(hx, hy);
  unionBox3(sdf, W, H, D, [cx - R, cy - R, cz - hz, cx + R, cy + R, cz + hz], (x, y, z) => {
    const qx = Math.abs((x - cx) * c + (y - cy) * s) - hx;
    const qy = Math.abs(-(x - cx) * s + (y - cy) * c) - hy;
    const qz = Math.abs(z - cz) - hz;
    return Math.hypot(Math.max(qx, 0), Math.max(qy, 0), Math.max(qz, 0)) + Math.min(Math.max(qx, qy, qz), 0);
  });
}

function addCylinderZ(sdf, W, H, D, cx, cy, r) {
  unionBox3(sdf, W, H, D, [cx - r, cy - r, 0, cx + r, cy + r, D - 1], (x, y) => Math.hypot(x - cx, y - cy) - r);
}

/** NACA0012 wing (chord, span, flat tips) centred at (cx, cy, cz), pitched nose-up by angleZ about mid-chord. */
function addWing(sdf, W, H, D, cx, cy, cz, chord, span, angleZ) {
  const a = angleZ || 0;
  const ox = cx - (chord / 2) * Math.cos(a);
  const oy = cy + (chord / 2) * Math.sin(a);
  const poly = nacaPolygon(0, 0, chord, 0.12, a, 200).map((p) => [p[0] + ox, p[1] + oy]);
  const b = bounds(poly);
  const section = new Float64Array(W * H).fill(NaN);
  unionBox3(sdf, W, H, D, [b[0], b[1], cz - span / 2, b[2], b[3], cz + span / 2], (x, y, z) => {
    const k = x + W * y;
    if (Number.isNaN(section[k])) section[k] = polygonDistance(poly, x, y);
    const d2 = section[k];
    const dz = Math.abs(z - cz) - span / 2;
    return d2 < 0 || dz < 0 ? Math.max(d2, dz) : Math.hypot(d2, dz);
  });
}

/** Frontal reference area (cells^2) for the force coefficients. */
function referenceArea(obstacle, size, D) {
  switch (obstacle) {
    case 'sphere': return (Math.PI * size * size) / 4;
    case 'cube': return size * size;
    case 'cylinder': return size * D;
    case 'wing': return size * 0.6 * D;
    default: return 0;
  }
}

// ---------------------------------------------------------------- units and analysis

/** tau = 3 nu + 1/2 with nu = u L / Re, clamped at 0.51 (src/units.ts:14-20). */
function deriveTau(re, uLat, length) {
  const raw = (3 * uLat * length) / re + 0.5;
  const tau = Math.max(raw, TAU_MIN);
  const nu = (tau - 0.5) / 3;
  return { nu, tau, clamped: raw < TAU_MIN, effectiveRe: (uLat * length) / nu, mach: uLat * Math.sqrt(3) };
}

function maxReynolds(uLat, length) {
  return (uLat * length * 3) / (TAU_MIN - 0.5);
}

function mean(v) {
  let s = 0;
  for (const x of v) s += x;
  return s / v.length;
}

/** streamFunction (src/analysis.ts:28-37): psi[y * W + x] from the velocity components. */
function streamFunction(ux, uy, W, H) {
  const psi = new Float64Array(W * H);
  for (let y = 1; y < H; y++) psi[y * W] = psi[(y - 1) * W] + 0.5 * (ux[y * W] + ux[(y - 1) * W]);
  for (let y = 0; y < H; y++) {
    for (let x = 1; x < W; x++) {
      const i = y * W + x;
      psi[i] = psi[i - 1] - 0.5 * (uy[i] + uy[i - 1]);
    }
  }
  return psi;
}

/** primaryVortex (src/analysis.ts:40-55): extremum of |psi| refined by a quadratic fit, [x, y, psi]. */
function primaryVortex(psi, W, H) {
  let best = 1;
  for (let y = 1; y < H - 1; y++) {
    for (let x = 1; x < W - 1; x++) if (Math.abs(psi[y * W + x]) > Math.abs(psi[best])) best = y * W + x;
  }
  const bx = best % W;
  const by = Math.floor(best / W);
  const refine = (m, c, p) => {
    const den = m - 2 * c + p;
    return den === 0 ? 0 : (0.5 * (m - p)) / den;
  };
  return [bx + refine(psi[best - 1], psi[best], psi[best + 1]), by + refine(psi[best - W], psi[best], psi[best + W]),
    psi[best]];
}

/** Period from linearly interpolated upward crossings of the mean; null with fewer than 3 crossings. */
function periodFromCrossings(t, v) {
  const m = mean(v);
  const times = [];
  for (let k = 1; k < v.length; k++) {
    const a = v[k - 1] - m;
    const b = v[k] - m;
    if (a < 0 && b >= 0) times.push(t[k - 1] + ((t[k] - t[k - 1]) * -a) / (b - a));
  }
  if (times.length < 3) return null;
  return { period: (times[times.length - 1] - times[0]) / (times.length - 1), crossings: times.length };
}

function coefficientScale(uTarget, area) {
  return area > 0 ? 2 / (uTarget * uTarget * area) : 0;
}

function hasDiverged(samples) {
  return samples.some((f) => !Number.isFinite(f.fx) || !Number.isFinite(f.fy) || (f.fz !== undefined && !Number.isFinite(f.fz)));
}

function fieldDiverged(values) {
  for (let k = 0; k < values.length; k++) if (!Number.isFinite(values[k])) return true;
  return false;
}

module.exports = {
  FAR, TAU_MIN, emptySdf, addCircle, addBox, nacaPolygon, addPolygon, polygonDistance, emptySdf3, addSphere, addBox3,
  streamFunction, primaryVortex,
  addCylinderZ, addWing, referenceArea, deriveTau, maxReynolds, mean, periodFromCrossings, coefficientScale,
  hasDiverged, fieldDiverged,
};
