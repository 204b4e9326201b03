'use strict';
// Headless frame driver: the step semantics of the reference's frame loop
// (web/js/main.js:383-415) without the UI.  One frame writes Params once
// (dt = (1/60) timeScale / substeps, collider velocity from the finite
// difference of its position, clamped to +-1.5 m/s) and submits the substeps
// in chunks of 6 command encoders.

const CHUNK = 6;

class FrameDriver {
  constructor(device, sim, cfg) {
    this.device = device;
    this.sim = sim;
    this.cfg = cfg;
    this.simTime = 0;
    this.input = { colliderPos: [0.5, 0.3, 0.5], colliderPrev: [0.5, 0.3, 0.5], colliderR: 0.08, pokeActive: false };
  }

  // dtWall: wall seconds since the previous frame (main.js:369-371 clamps it to 1/30)
  frame(dtWall) {
    const cfg = this.cfg, input = this.input;
    dtWall = Math.min(dtWall === undefined ? 1 / 60 : dtWall, 1 / 30);
    const substeps = cfg.substeps;
    const dt = (1 / 60) * cfg.timeScale / substeps;
    const cv = [0, 1, 2].map(i => (input.colliderPos[i] - input.colliderPrev[i]) / Math.max(dtWall, 1e-3));
    input.colliderPrev = input.colliderPos.slice();
    this.sim.writeParams({
      dt, domainMin: cfg.domainMin, domainMax: cfg.domainMax, gravity: cfg.gravity,
      colliderPos: input.colliderPos, colliderR: input.colliderR,
      colliderVel: cv.map(v => Math.max(-1.5, Math.min(1.5, v || 0))),
      colliderActive: input.pokeActive,
      xsphC: cfg.xsphC, omega: cfg.omega, contactOmega: cfg.contactOmega,
      pbfEps: cfg.pbfEps, sCorrK: cfg.sCorrK, time: this.simTime,
      kDamp: cfg.kDamp === undefined ? 0.4 : cfg.kDamp,
    });
    for (let s = 0; s < substeps; s += CHUNK) {
      const enc = this.device.createCommandEncoder();
      for (let k = s; k < Math.min(s + CHUNK, substeps); k++) this.sim.encodeSubstep(enc, cfg.iterations, cfg.hasWater);
      this.device.queue.submit([enc.finish()]);
    }
    this.simTime += (1 / 60) * cfg.timeScale;
  }
}

module.exports = { FrameDriver, CHUNK };
