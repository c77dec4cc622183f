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

// The gentle pour from a pot with a rim lip in place of the tube spout (build-only geometry).
function lipPour3D(cellMm) {
  return load().lipPour(Number(cellMm));
}

// The 6.2 x 14 cm pot with a 2 cm tube spout 2 cm above the floor, ramped to 90 deg, CFL 1
// (build-only preset).
function lowSpoutPour3D(cellMm, rateDegS = 9) {
  return load().lowSpoutPour(Number(cellMm), Number(rateDegS));
}

module.exports = { gentlePour3D, rampPour3D, lipPour3D, lowSpoutPour3D };
