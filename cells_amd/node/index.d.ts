// Type declarations for cells_amd/node (the drop-in for the reference's src/solver.ts and src/solver3d.ts).

export declare const enum XModeValues { Periodic = 0, Wall = 1, Open = 2 }
export declare const XMode: { readonly Periodic: 0; readonly Wall: 1; readonly Open: 2 };
export declare const YMode: { readonly Periodic: 0; readonly NoSlip: 1; readonly Slip: 2; readonly MovingWall: 3 };
export type Precision = 'fp16' | 'fp32' | 'bf16' | 'fp16sr';
export declare const MAGIC_LAMBDA: number;
export declare const HISTORY_LEN: number;
export declare const PRESETS3: { readonly low: [number, number, number]; readonly medium: [number, number, number]; readonly high: [number, number, number] };

export interface SolverConfig {
  width: number; height: number;
  left: number; right: number; bottom: number; top: number;
  tau: number;
  lambda?: number;
  smagorinsky?: number;
  uIn?: number;
  parabolic?: boolean;
  uLid?: number;
  spongeFraction?: number;
  spongeTau?: number;
  bodyForce?: [number, number];
  absorb?: number;
  inletLayerFraction?: number;
  workgroupSize?: number;
  forceEvery?: number;
}

export interface Solver3DConfig {
  width: number; height: number; depth: number;
  precision: Precision;
  tau: number;
  lambda?: number;
  smagorinsky?: number;
  uIn?: number;
  spongeFraction?: number;
  spongeTau?: number;
  absorb?: number;
  workgroupSize?: number;
  forceEvery?: number;
  /** 19 (default, the reference's D3Q19) or 27 (D3Q27). */
  velocitySet?: 19 | 27;
}

export interface ForceSample { step: number; fx: number; fy: number }
export interface ForceSample3 { step: number; fx: number; fy: number; fz: number }
export type RampKind = 'off' | 'linear' | 'smoothstep';

declare class SolverBase {
  readonly step: number;
  setSdf(sdf: Float32Array): void;
  resetForces(): void;
  setInletRamp(kind: RampKind, uTarget: number, steps: number): void;
  encodeSteps(pass: unknown, n: number): void;
  encodeMacro(pass?: unknown): void;
  run(n: number, chunk?: number): Promise<void>;
  readMacro(): Promise<Float32Array>;
  readPopulations(): Float32Array | Uint16Array;
  readFlags(): Uint32Array;
  timeSteps(n: number): number;
  autotune(trialSteps?: number): { block: number; sweep: number; tiled: boolean; stride: number };
  /** Device population layout by hand (tiles of 64 cells, or a padded stride; 0 = unpadded). */
  setPopulationLayout(tiled: boolean, stride?: number): void;
  /** First step whose force sample was non-finite since initField / resetForces, 0 while none was. */
  divergedAt(): number;
  destroy(): void;
}

export declare class Solver extends SolverBase {
  static create(device: number | unknown, config: SolverConfig): Promise<Solver>;
  readonly W: number; readonly H: number; readonly N: number;
  readonly cfg: Required<SolverConfig>;
  update(changes: Partial<SolverConfig>): void;
  initField(macro?: Float32Array): void;
  rebuildFlags(): void;
  brush(cx: number, cy: number, r: number, erase: boolean, rebuild?: boolean): void;
  brushStroke(points: [number, number][], r: number, erase: boolean): void;
  readForces(): Promise<ForceSample[]>;
}

export declare class Solver3D extends SolverBase {
  static create(device: number | unknown, config: Solver3DConfig): Promise<Solver3D>;
  readonly W: number; readonly H: number; readonly D: number; readonly N: number;
  readonly cfg: Solver3DConfig;
  setTau(tau: number): void;
  setSmagorinsky(cs: number): void;
  setInlet(u: number): void;
  /** Open x planes: the reference's equilibrium planes (default) or Hecht-Harting Zou-He faces. */
  setOpenFaces(kind: 'equilibrium' | 'zouHe'): void;
  initField(): void;
  readForces(): Promise<ForceSample3[]>;
}

export interface ShapeAdjointOptions {
  steps: number;
  window: number;
  direction?: number[];
  scale?: number;
  checkpointEvery?: number;
  ramp?: { kind: RampKind | 'none'; uTarget: number; steps: number } | null;
  recordBudgetBytes?: number;
}

export declare class ShapeAdjoint {
  constructor(solver: Solver | Solver3D, opts: ShapeAdjointOptions);
  forward(sdf?: Float32Array): Promise<number>;
  backward(seed?: number): Promise<Float64Array>;
}

export declare function trtLambda3(tau: number): number;
export declare function bytesPerPopulation(p: Precision): number;
export declare function fitsLimits(W: number, H: number, D: number, p: Precision, limits: { maxStorageBufferBindingSize: number; maxBufferSize: number }): boolean;

export declare const FAR: number;
export declare const TAU_MIN: number;
export declare function emptySdf(W: number, H: number): Float32Array;
export declare function addCircle(sdf: Float32Array, W: number, H: number, cx: number, cy: number, r: number): void;
export declare function addBox(sdf: Float32Array, W: number, H: number, cx: number, cy: number, hx: number, hy: number, angle?: number): void;
export declare function nacaPolygon(x0: number, y0: number, chord: number, thickness: number, alpha: number, n?: number): [number, number][];
export declare function addPolygon(sdf: Float32Array, W: number, H: number, poly: [number, number][]): void;
export declare function emptySdf3(W: number, H: number, D: number): Float32Array;
export declare function addSphere(sdf: Float32Array, W: number, H: number, D: number, cx: number, cy: number, cz: number, r: number): void;
export declare function addBox3(sdf: Float32Array, W: number, H: number, D: number, cx: number, cy: number, cz: number, hx: number, hy: number, hz: number, angleZ?: number): void;
export declare function addCylinderZ(sdf: Float32Array, W: number, H: number, D: number, cx: number, cy: number, r: number): void;
export declare function addWing(sdf: Float32Array, W: number, H: number, D: number, cx: number, cy: number, cz: number, chord: number, span: number, angleZ?: number): void;
export declare function referenceArea(obstacle: 'sphere' | 'cube' | 'cylinder' | 'wing' | 'none', size: number, D: number): number;
export declare function deriveTau(re: number, uLat: number, length: number): { nu: number; tau: number; clamped: boolean; effectiveRe: number; mach: number };
export declare function maxReynolds(uLat: number, length: number): number;
export declare function mean(v: number[]): number;
export declare function streamFunction(ux: ArrayLike<number>, uy: ArrayLike<number>, W: number, H: number): Float64Array;
export declare function primaryVortex(psi: ArrayLike<number>, W: number, H: number): [number, number, number];
export declare function periodFromCrossings(t: number[], v: number[]): { period: number; crossings: number } | null;
export declare function coefficientScale(uTarget: number, area: number): number;
export declare function hasDiverged(samples: ReadonlyArray<{ fx: number; fy: number; fz?: number }>): boolean;
export declare function fieldDiverged(values: Float32Array): boolean;
