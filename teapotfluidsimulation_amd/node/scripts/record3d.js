'use strict';
// Record the offline 3D pour for browser playback, as scripts/record3d.ts does with the CPU
// solver: every `stride`-th parcel, 10 frames per simulated second, gzip level 9.  The frames are
// quantised and delta-encoded on the GPU (Recorder3D); this script keeps the reference's frame
// schedule and header.
// Usage: node scripts/record3d.js [cellMm] [stride] [out]
const fs = require('fs');
const zlib = require('zlib');
const { performance } = require('perf_hooks');
const { gentlePour3D } = require('../src/sim3d/presets3d');
const { Simulation3D, Recorder3D } = require('../src/sim3d/solver3d');
const { writeRecording } = require('../src/view3d/format');

const [cellArg = '5', strideArg = '3', out = 'pour3d.bin.gz'] = process.argv.slice(2);
const stride = Number(strideArg);
const fps = 10;
const quantum = 0.00025;
const cfg = gentlePour3D(Number(cellArg), 14);
const sim = new Simulation3D(cfg);
const rec = new Recorder3D(sim, stride, quantum);
const frames = [];
const angles = [];
const poured = [];
const perParcelMl = sim.parcelVolume * 2 * 1e6;
const capture = () => {
  frames.push(rec.capture());
  angles.push(sim.diagnostics().angleDeg);
  poured.push(sim.stats.dischargedByOrigin.map((n) => n * perParcelMl));
};
const t0 = performance.now();
let doneAt = -1;
capture();
while (!sim.finished) {
  const d = sim.advance(1 / fps);
  capture();
  if (d.phase === 'done' && doneAt < 0) doneAt = d.time;
  if (doneAt >= 0 && d.time > doneAt + 1.5) break;
  if (frames.length % 10 === 0)
    process.stderr.write(
      `t=${d.time.toFixed(1)} poured=${d.dischargedMl.toFixed(0)} mL ${((performance.now() - t0) / 1000).toFixed(0)}s\n`,
    );
}
const shares = sim.dischargedFractions() || [0, 0, 0];
const header = {
  version: 1,
  frames: frames.length,
  parcels: rec.parcels,
  fps,
  quantum,
  cellMm: Number(cellArg),
  simulatedParcels: sim.n,
  teapot: cfg.teapot,
  pivot: cfg.pivot,
  pivotWorld: sim.pivotWorld,
  zOffset: sim.h,
  domain: [sim.nx * sim.h, sim.ny * sim.h, sim.nz * sim.h],
  angles,
  poured,
  finalShares: shares,
  label: `Recorded offline 3D simulation (${Number(cellArg)} mm cells, ${sim.n} parcels in the half pot); every ${stride}th parcel shown.`,
};
const raw = writeRecording(header, rec.origin, frames);
rec.destroy();
sim.destroy();
const gz = zlib.gzipSync(raw, { level: 9 });
fs.writeFileSync(out, gz);
console.log(
  `wrote ${out}: ${frames.length} frames, ${header.parcels} parcels, ${(raw.length / 1e6).toFixed(1)} MB raw, ${(gz.length / 1e6).toFixed(1)} MB gzip; poured ${poured[poured.length - 1].reduce((a, b) => a + b, 0).toFixed(1)} mL, shares ${shares.map((s) => (s * 100).toFixed(1)).join('/')} %`,
);
