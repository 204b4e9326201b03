'use strict';
// Headless frame driver: the step semantics of the reference's frame loop
// (web/js/main.js:383-415) without the UI.  One frame writes Params once
// (dt = (1/60) timeScale / substeps, collider velocity from the finite
// difference of its position, clamped to +-1.5 m/s) and submits the substeps
// in chunks of 6 command encoders.

const CHUNK = 6;

// the Params of one frame (main.js:386-402): dt, per-frame collider velocity by
// finite difference clamped to +-1.5 m/s; input.colliderPrev advances
function frameParams(cfg, input, dtWall, simTime) {
  const dt = cfg.dt || (1 / 60) * cfg.timeScale / cfg.substeps;
  const cv = [0, 1, 2].map(i => (input.colliderPos[i] - input.colliderPrev[i]) / Math.max(dtWall, 1e-3));
  input.colliderPrev = input.colliderPos.slice();
  return {
    dt, domainMin: cfg.domainMin, domainMax: cfg.domainMax, gravity: cfg.gravity,
    colliderPos: input.colliderPos, colliderR: input.colliderR,
    colliderVel: cv.map(v => Math.max(-1.5, Math.min(1.5, v || 0))),
    colliderActive: input.pokeActive,
    xsphC: cfg.xsphC, omega: cfg.omega, contactOmega: cfg.contactOmega,
    pbfEps: cfg.pbfEps, sCorrK: cfg.sCorrK, time: simTime,
    kDamp: cfg.kDamp === undefined ? 0.4 : cfg.kDamp,
  };
}

function defaultInput() {
  return { colliderPos: [0.5, 0.3, 0.5], colliderPrev: [0.5, 0.3, 0.5], colliderR: 0.08, pokeActive: false };
}

class FrameDriver {
  // opts.graph: replay each chunk from the simulation's captured HIP graph
  // (xpbi_graph_step; same results) instead of command encoders
  constructor(device, sim, cfg, opts) {
    this.device = device;
    this.sim = sim;
    this.cfg = cfg;
    this.graph = !!(opts && opts.graph);
    this.simTime = 0;
    this.input = defaultInput();
  }

  // dtWall: wall seconds since the previous frame (main.js:369-371 clamps it to 1/30)
  frame(dtWall) {
    const cfg = this.cfg;
    dtWall = Math.min(dtWall === undefined ? 1 / 60 : dtWall, 1 / 30);
    const substeps = cfg.substeps;
    this.sim.writeParams(frameParams(cfg, this.input, dtWall, this.simTime));
    for (let s = 0; s < substeps; s += CHUNK) {
      const n = Math.min(CHUNK, substeps - s);
      if (this.graph) {
        this.sim.graphStep(n, cfg.iterations, cfg.hasWater);
        continue;
      }
      const enc = this.device.createCommandEncoder();
      for (let k = 0; k < n; k++) this.sim.encodeSubstep(enc, cfg.iterations, cfg.hasWater);
      this.device.queue.submit([enc.finish()]);
    }
    this.simTime += (1 / 60) * cfg.timeScale;
  }
}

module.exports = { FrameDriver, CHUNK, frameParams, defaultInput };
