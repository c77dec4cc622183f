'use strict';
// Simulation3D and Pcg3D with the API of src/sim3d/solver3d.ts, running on the GPU through
// libpourlab3d.so (N-API binding).  Method names, argument meaning and returned shapes follow the
// reference, so a driver written like benchmarks/run3d.ts runs unchanged:
//
//   const sim = new Simulation3D(gentlePour3D(4, 20));
//   while (!sim.finished) { const d = sim.advance(0.5); ... }
//
// The GpuSimulation3D extras (src/gpu3d/solver.ts: stats, readParticles, readRenderPositions,
// controller.setTarget, destroy) are methods of the same class; every call blocks until the
// device work it issued is done.
const { load } = require('../addon');

// Parcel status (solver3d.ts:20-23)
const IN_POT = 0;
const DISCHARGED = 1;
const SPILLED = 2;
const LOST = 3;
// Cell classification (solver3d.ts:313-336)
const AIR = 0;
const FLUID = 1;
const SOLID = 2;

// STAT slots (gpu3d/layout.ts:20-37)
const STAT = {
  inPot: 0,
  discharged: 1,
  spilled: 2,
  lost: 3,
  originBottom: 4,
  originMiddle: 5,
  originTop: 6,
  pressureFailures: 7,
  maxSpeedSq: 8,
  maxDivergence: 9,
  interiorCount: 10,
  interiorSum: 11,
  pressureIterations: 12,
  correctionIterations: 13,
};
const SCALE_RATIO = 1 << 16;

class Simulation3D {
  constructor(cfg) {
    const lib = load();
    this.cfg = cfg;
    this._lib = lib;
    this._h = lib.create(cfg);
    const info = lib.info(this._h);
    this.n = info.n;
    this.nx = info.nx;
    this.ny = info.ny;
    this.nz = info.nz;
    this.h = info.h;
    this.fillHeight = info.fillHeight;
    this.pivotWorld = info.pivotWorld;
    this.parcelVolume = info.parcelVolume;
    this.bandCounts = info.bandCounts;
    this.maxLiquidCells = info.maxLiquidCells;
    const self = this;
    // PourController surface used by the live view (motion.ts:40-44)
    this.controller = {
      setTarget(angleDeg) {
        self._lib.setTarget(self._h, angleDeg);
      },
    };
  }

  get finished() {
    return this._lib.finished(this._h);
  }

  get time() {
    return this.diagnostics().time;
  }

  get steps() {
    return this.diagnostics().steps;
  }

  stableDt() {
    return this._lib.stableDt(this._h);
  }

  step(dt) {
    this._lib.step(this._h, dt);
  }

  advance(seconds) {
    return this._lib.advance(this._h, seconds);
  }

  diagnostics() {
    return this._lib.diagnostics(this._h);
  }

  dischargedFractions() {
    return this._lib.dischargedFractions(this._h);
  }

  // Public parcel fields (solver3d.ts:75-85), read back from the device on access.
  state() {
    return this._lib.readState(this._h, this.n);
  }

  get x() {
    return this.state().x;
  }

  get y() {
    return this.state().y;
  }

  get z() {
    return this.state().z;
  }

  get vx() {
    return this.state().vx;
  }

  get vy() {
    return this.state().vy;
  }

  get vz() {
    return this.state().vz;
  }

  get status() {
    return this.state().status;
  }

  get origin() {
    return this.state().origin;
  }

  get dischargeOrder() {
    return Array.from(this._lib.dischargeOrder(this._h));
  }

  // GpuSimulation3D surface (gpu3d/solver.ts)
  get stats() {
    const raw = this._lib.stats(this._h);
    const f = new Float32Array(raw.buffer);
    const ic = raw[STAT.interiorCount];
    return {
      counts: [raw[0], raw[1], raw[2], raw[3]],
      dischargedByOrigin: [raw[4], raw[5], raw[6]],
      pressureFailures: raw[STAT.pressureFailures],
      maxSpeedSq: f[STAT.maxSpeedSq],
      maxDivergence: f[STAT.maxDivergence],
      interiorDensity: ic > 0 ? raw[STAT.interiorSum] / SCALE_RATIO / ic : 1,
      pressureIterations: raw[STAT.pressureIterations],
      correctionIterations: raw[STAT.correctionIterations],
    };
  }

  // APIC affine rows, Float64Array(9n): [9p + 3c + d] = d(velocity c)/d(axis d)
  affine() {
    return this._lib.readAffine(this._h, this.n);
  }

  readParticles() {
    return this._lib.readParticles(this._h, this.n);
  }

  readRenderPositions() {
    return this._lib.readRenderPositions(this._h, this.n);
  }

  sync() {
    // calls block until their device work is done; nothing is in flight between calls
  }

  destroy() {
    if (this._h) this._lib.destroy(this._h);
    this._h = null;
  }
}

class Recorder3D {
  // Device-side frame encoder of a recording (capture/encode of scripts/record3d.ts:23-36 and
  // encodeRecording, format.ts:46-76): every `stride`-th parcel, quantised to `quantum` metres.
  // capture() returns the next frame's Uint16Array(3P): u16 values for the first frame, int16
  // differences after it.  The simulation cannot be destroyed while its recorder is live.
  constructor(sim, stride = 3, quantum = 0.00025) {
    this._lib = load();
    this.stride = stride;
    this.quantum = quantum;
    this._r = this._lib.recorderCreate(sim._h, stride, quantum);
    const p = this._lib.recorderParcels(this._r);
    this.parcels = p.count;
    this.origin = p.origin;
  }

  capture() {
    return this._lib.recorderCapture(this._r, this.parcels);
  }

  destroy() {
    if (this._r) this._lib.recorderDestroy(this._r);
    this._r = null;
  }
}

class Pcg3D {
  // Pcg3D (solver3d.ts:827-974): matrix-free 7-point PCG on the FLUID cells, on the GPU.
  constructor(nx, ny, nz, preconditioner = 'mic0') {
    this.nx = nx;
    this.ny = ny;
    this.nz = nz;
    this.preconditioner = preconditioner;
  }

  // Fills q and returns {iterations, converged, residual, trueResidual, liquidCells, kernelMs}.
  solve(cell, b, q, tol, maxIter) {
    return load().pcgSolve(this.nx, this.ny, this.nz, cell, b, q, tol, maxIter, this.preconditioner);
  }
}

module.exports = { Simulation3D, Recorder3D, Pcg3D, IN_POT, DISCHARGED, SPILLED, LOST, AIR, FLUID, SOLID, STAT };
