'use strict';
// GpuSimulation3D with the asynchronous API of src/gpu3d/solver.ts:89-490, so the live view
// (src/view3d/live.ts, its LiveSim interface at :29-53) can drive the HIP solver unchanged:
//
//   const sim = new GpuSimulation3D(null, liveConfig('gentle-5'));   // the device argument is unused
//   while (!sim.finished) { await sim.step(); const xyz = await sim.readRenderPositions(); }
//
// Calls run on the GPU through the N-API binding and block until their device work is done; the
// Promises resolve with the results.  The controller, stableDt and the stats are computed on the
// device from the substep just completed, so there is no one-substep-late stats pipeline and
// sync() has nothing to flush.
const { Simulation3D } = require('../sim3d/solver3d');

class GpuSimulation3D {
  constructor(device, cfg) {
    this._sim = new Simulation3D(cfg);
    this.cfg = cfg;
    const sim = this._sim;
    const self = this;
    // PourController surface (motion.ts:40-44, :83-94)
    this.controller = {
      get angleDeg() {
        return sim.diagnostics().angleDeg;
      },
      get phase() {
        return sim.diagnostics().phase;
      },
      setTarget(angleDeg) {
        sim.controller.setTarget(angleDeg);
      },
    };
    // the CPU-side facts the view reads (live.ts:39-48)
    this.cpu = {
      get parcelVolume() {
        return sim.parcelVolume;
      },
      get origin() {
        return sim.origin;
      },
      get x() {
        return sim.x;
      },
      get y() {
        return sim.y;
      },
      get z() {
        return sim.z;
      },
      get h() {
        return sim.h;
      },
      get pivotWorld() {
        return sim.pivotWorld;
      },
      get cfg() {
        return self.cfg;
      },
    };
  }

  get n() {
    return this._sim.n;
  }

  get time() {
    return this._sim.time;
  }

  get steps() {
    return this._sim.steps;
  }

  get finished() {
    return this._sim.finished;
  }

  get stats() {
    return this._sim.stats;
  }

  stableDt() {
    return this._sim.stableDt();
  }

  async step(dt) {
    const d = dt === undefined ? Math.min(this._sim.stableDt(), this.cfg.duration - this._sim.time) : dt;
    this._sim.step(d);
  }

  async advance(duration) {
    return this._sim.advance(duration);
  }

  async sync() {}

  async readParticles() {
    return this._sim.readParticles();
  }

  async readRenderPositions() {
    return this._sim.readRenderPositions();
  }

  destroy() {
    this._sim.destroy();
  }
}

module.exports = { GpuSimulation3D };
