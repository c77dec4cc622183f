'use strict';
// Verify the GPU pour against the CPU solver's published results, with the criteria of
// scripts/verify-gpu.ts: poured volume within 1 %, layer shares within 2 points, no lost parcels.
// The reference runs its WebGPU port in a browser; here the same pours run through the N-API
// binding in this process.
// Usage: node scripts/verify-gpu.js [cellMm ...] [--cpu path/to/run3d.json] > gpu3d.json
const fs = require('fs');
const { performance } = require('perf_hooks');
const { gentlePour3D } = require('../src/sim3d/presets3d');
const { Simulation3D } = require('../src/sim3d/solver3d');

// CPU rows of benchmarks/run3d.json (:6-20, :27-41): FP64 TypeScript solver, seed 1.
const PUBLISHED_CPU_ROWS = [
  {
    cellMm: 5,
    dischargedMl: 406.7796610169492,
    lostMl: 0,
    dischargedFractions: [0.2770669291338583, 0.5194389763779528, 0.20349409448818898],
    simulatedSeconds: 11.547468892151917,
    wallSeconds: 683.50239039,
  },
  {
    cellMm: 4,
    dischargedMl: 410.53502295849466,
    lostMl: 0,
    dischargedFractions: [0.2921995147826802, 0.5096426662280521, 0.19815781898926765],
    simulatedSeconds: 11.543841673447004,
    wallSeconds: 1636.373419337,
  },
];

function pour(cellMm, preconditioner) {
  const cfg = gentlePour3D(cellMm, 20);
  if (preconditioner) cfg.preconditioner = preconditioner;
  const sim = new Simulation3D(cfg);
  const t0 = performance.now();
  let doneAt = -1;
  let maxIt = 0;
  while (!sim.finished) {
    const d = sim.advance(0.5);
    maxIt = Math.max(maxIt, d.pressureIterations);
    if (d.phase === 'done' && doneAt < 0) doneAt = d.time;
    if (doneAt >= 0 && d.time > doneAt + 1.5) break;
  }
  const wall = (performance.now() - t0) / 1000;
  const d = sim.diagnostics();
  const row = {
    cellMm,
    parcelsHalfPot: sim.n,
    grid: [sim.nx, sim.ny, sim.nz],
    simulatedSeconds: d.time,
    steps: d.steps,
    dischargedMl: d.dischargedMl,
    spilledMl: d.spilledMl,
    lostMl: d.lostMl,
    dischargedFractions: sim.dischargedFractions(),
    pressureFailures: d.pressureFailures,
    lastPressureIterations: maxIt,
    interiorDensity: d.interiorDensity,
    wallSeconds: wall,
    realTimeFactor: d.time / wall,
  };
  sim.destroy();
  return row;
}

const argv = process.argv.slice(2);
let cpuRows = PUBLISHED_CPU_ROWS;
let preconditioner = null;
const cells = [];
for (let i = 0; i < argv.length; i++) {
  if (argv[i] === '--cpu') cpuRows = JSON.parse(fs.readFileSync(argv[++i], 'utf8')).rows;
  else if (argv[i] === '--preconditioner') preconditioner = argv[++i];
  else cells.push(Number(argv[i]));
}
if (cells.length === 0) cells.push(5);
const rows = [];
let allPass = true;
for (const cellMm of cells) {
  const r = pour(cellMm, preconditioner);
  const cpu = cpuRows.find((c) => c.cellMm === cellMm);
  const check = cpu && {
    cpuDischargedMl: cpu.dischargedMl,
    cpuFractions: cpu.dischargedFractions,
    cpuRealTimeFactor: cpu.simulatedSeconds / cpu.wallSeconds,
    volumeDifferencePercent: (100 * (r.dischargedMl - cpu.dischargedMl)) / cpu.dischargedMl,
    shareDifferencePoints: r.dischargedFractions.map((f, k) => 100 * (f - cpu.dischargedFractions[k])),
  };
  const pass =
    !!check &&
    Math.abs(check.volumeDifferencePercent) <= 1 &&
    check.shareDifferencePoints.every((x) => Math.abs(x) <= 2) &&
    r.lostMl === 0;
  allPass = allPass && pass;
  rows.push(Object.assign({}, r, { comparison: check || null, pass }));
  process.stderr.write(`${cellMm} mm: ${JSON.stringify(Object.assign({}, check, { rtf: r.realTimeFactor, pass }))}\n`);
}
console.log(
  JSON.stringify(
    {
      generatedAt: new Date().toISOString(),
      device: 'MI355X (gfx950) via libpourlab3d.so',
      criteria: 'poured volume within 1 %, layer shares within 2 points, no lost parcels (scripts/verify-gpu.ts)',
      rows,
    },
    null,
    1,
  ),
);
process.exitCode = allPass ? 0 : 1;
