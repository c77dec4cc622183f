'use strict';
// gentlePour3D (src/sim3d/presets3d.ts:9-26): the gentle 400 mL pour preset as a Config3D object.
// The values come from the library (pl3d_gentle_pour), so the JS, Python and C hosts share one
// definition.
const { load } = require('../addon');

function gentlePour3D(cellMm, seconds) {
  return load().gentlePour(Number(cellMm), Number(seconds));
}

// The gentle pour with a constant-rate ramp to maxAngleDeg (build-only motion kind 'ramp').
function rampPour3D(cellMm, rateDegS, maxAngleDeg = 90) {
  const cfg = gentlePour3D(cellMm, 20);
  cfg.motion = Object.assign({}, cfg.motion, {
    kind: 'ramp',
    holdTiltRate: Number(rateDegS),
    maxAngleDeg: Number(maxAngleDeg),
  });
  return cfg;
}

module.exports = { gentlePour3D, rampPour3D };
