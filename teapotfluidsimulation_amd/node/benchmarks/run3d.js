'use strict';
// The offline 3D run of the gentle 400 mL pour, written as benchmarks/run3d.ts drives the
// reference, on the GPU.  Usage: node benchmarks/run3d.js [cellSizeMm] [seconds] > run3d.json
const { performance } = require('perf_hooks');
const { gentlePour3D } = require('../src/sim3d/presets3d');
const { Simulation3D } = require('../src/sim3d/solver3d');

const [cellArg = '4', secondsArg = '20'] = process.argv.slice(2);
const sim = new Simulation3D(gentlePour3D(Number(cellArg), Number(secondsArg)));
const t0 = performance.now();
const rows = [];
process.stderr.write(`3D: ${sim.n} parcels (half pot), grid ${sim.nx}x${sim.ny}x${sim.nz}\n`);
let doneAt = -1;
while (!sim.finished) {
  const d = sim.advance(0.5);
  rows.push(d);
  process.stderr.write(
    `t=${d.time.toFixed(1)} ${d.phase} tilt=${d.angleDeg.toFixed(1)} out=${d.dischargedMl.toFixed(0)} spill=${d.spilledMl.toFixed(0)} lost=${d.lostMl.toFixed(0)} it=${d.pressureIterations} rho=${d.interiorDensity.toFixed(3)} wall=${((performance.now() - t0) / 1000).toFixed(1)}s\n`,
  );
  if (d.phase === 'done' && doneAt < 0) doneAt = d.time;
  if (doneAt >= 0 && d.time > doneAt + 1.5) break;
}
const d = sim.diagnostics();
console.log(
  JSON.stringify(
    {
      generatedAt: new Date().toISOString(),
      cellMm: Number(cellArg),
      result: d,
      dischargedFractions: sim.dischargedFractions(),
      wallSeconds: (performance.now() - t0) / 1000,
      history: rows,
    },
    null,
    1,
  ),
);
sim.destroy();
