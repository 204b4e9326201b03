'use strict';
// XPBI simulation host: the surface of the reference's web/js/sim.js (buffer
// set-up, writeMaterials, writeParams, encodeSubstep, profileSubstep) over the
// MI355X substep engine.  The compute kernels are compiled into libxpbi.so and
// reached through the N-API addon; the GPU "device" here is a HIP device
// ordinal, and a command encoder records substeps that queue.submit() runs.
// CommonJS + ES2019 syntax (the Node in this image is v12).

const path = require('path');

let addon = null;
function loadAddon() {
  if (!addon) addon = require(path.join(__dirname, 'xpbi_napi.node'));
  return addon;
}

const MAT_STRIDE = 16; // floats per material (sim.js:4)
const MAT = { ELASTIC: 0, DP: 1, NACC: 2, VM: 3, HB: 4, SNOW: 5, WATER: 6 };
const FIELD = { POS: 0, VEL: 1, F: 2, L: 3, AUX: 4, SP: 5, CELLSTART: 6, PERM: 7, ID: 8, MATS: 9 };

function def(v, d) { return v === undefined || v === null ? d : v; }

// ---------------------------------------------------------------- device shim
// requestDevice({ordinal}) -> { queue.submit(commandBuffers), createCommandEncoder() }
class CommandEncoder {
  constructor() { this.cmds = []; }
  finish() { return { cmds: this.cmds }; }
}
function requestDevice(opts) {
  const ordinal = opts && opts.ordinal ? opts.ordinal : 0;
  return {
    ordinal,
    createCommandEncoder() { return new CommandEncoder(); },
    queue: {
      // consecutive substeps of one simulation are issued as one xpbi_step call
      submit(buffers) {
        for (const b of buffers) {
          let k = 0;
          while (k < b.cmds.length) {
            const c = b.cmds[k];
            let n = 1;
            while (k + n < b.cmds.length && b.cmds[k + n].sim === c.sim &&
                   b.cmds[k + n].iterations === c.iterations && b.cmds[k + n].hasWater === c.hasWater) n++;
            loadAddon().step(c.sim.handle, n, c.iterations, !!c.hasWater);
            k += n;
          }
        }
      },
      onSubmittedWorkDone() { return Promise.resolve(); },
    },
  };
}

// grid of one scene: sized to its domain with a 2-cell pad (sim.js:20-26)
function sceneGrid(scene) {
  const cs = 2 * scene.h;
  const pad = 2;
  const gridOrigin = scene.domainMin.map(v => v - pad * cs);
  const gridDims = [0, 1, 2].map(i => Math.ceil((scene.domainMax[i] - scene.domainMin[i]) / cs) + 2 * pad);
  return { h: scene.h, support: cs, V0: scene.h * scene.h * scene.h, gridOrigin, gridDims,
           numCells: gridDims[0] * gridDims[1] * gridDims[2], N: scene.positions.length / 3 };
}

class Simulation {
  // scene: one scene (the reference's constructor, sim.js:10-30) or an array of
  // scenes, run as one batch: one launch per kernel for all of them, each with
  // its own grid, materials and Params (writeMaterials / writeParams take the
  // scene's index)
  constructor(device, scene) {
    this.device = device;
    this.scenes = Array.isArray(scene) ? scene : [scene];
    this.scene = this.scenes[0];
    this.grids = this.scenes.map(sceneGrid);
    const g = this.grids[0];
    this.N = this.grids.reduce((a, q) => a + q.N, 0);
    this.h = g.h;
    this.support = g.support;
    this.V0 = g.V0;
    this.frame = 0;
    this.gridOrigin = g.gridOrigin;
    this.gridDims = g.gridDims;
    this.numCells = g.numCells;
    this.createBuffers();
    this.createPipelines();
  }

  // sim.js:32-82: the library allocates every device buffer and uploads the state
  createBuffers() {
    const f32 = a => (a instanceof Float32Array ? a : Float32Array.from(a));
    const descs = this.scenes.map((s, q) => {
      const N = this.grids[q].N;
      let tags = s.tags ? f32(s.tags) : null;
      if (!tags) {
        tags = new Float32Array(N);
        for (let i = 0; i < N; i++) tags[i] = s.matIds[i];
      }
      return {
        N, positions: f32(s.positions), tags,
        states: s.states ? f32(s.states) : undefined,
        velocities: s.velocities ? f32(s.velocities) : undefined,
        h: s.h, domainMin: s.domainMin, domainMax: s.domainMax,
        device: this.device ? this.device.ordinal : 0,
      };
    });
    this.handle = descs.length === 1 ? loadAddon().create(descs[0]) : loadAddon().createBatch(descs);
    this.grids.forEach((g, q) => {
      const info = loadAddon().sceneInfo(this.handle, q);
      if (info.numCells !== g.numCells || info.gridDims.some((d, i) => d !== g.gridDims[i])) {
        throw new Error('grid mismatch between the host and libxpbi');
      }
      g.firstParticle = info.firstParticle;
      g.firstCell = info.firstCell;
    });
    this.totalCells = this.grids.reduce((a, q) => a + q.numCells, 0);
    this.scenes.forEach((s, q) => this.writeMaterials(s.materials, q));
  }

  // sim.js:84-99 (scene: index in a batch)
  writeMaterials(materials, scene) {
    const q = scene || 0;
    const V0 = this.grids[q].V0;
    const arr = new Float32Array(8 * MAT_STRIDE);
    const u32 = new Uint32Array(arr.buffer);
    materials.forEach((m, i) => {
      const o = i * MAT_STRIDE;
      const mu = m.E / (2 * (1 + m.nu));
      const lam = m.E * m.nu / ((1 + m.nu) * (1 - 2 * m.nu));
      arr[o + 0] = mu; arr[o + 1] = lam; arr[o + 2] = m.density;
      u32[o + 3] = m.kind;
      arr[o + 4] = def(m.p0, 0); arr[o + 5] = def(m.p1, 0); arr[o + 6] = def(m.p2, 0); arr[o + 7] = def(m.p3, 0);
      const c = m.color || [0, 0, 0];
      arr[o + 8] = c[0]; arr[o + 9] = c[1]; arr[o + 10] = c[2]; arr[o + 11] = 1;
      arr[o + 12] = def(m.boundFriction, 0.4);
      arr[o + 13] = m.density * V0;
    });
    loadAddon().writeSceneMaterials(this.handle, q, new Uint8Array(arr.buffer, 0, 64 * Math.min(8, materials.length)));
  }

  // the kernels are compiled once into libxpbi.so (sim.js:101-156 built WGSL pipelines)
  createPipelines() {}

  // sim.js:159-186: the 256-B Params record (scene: index in a batch)
  writeParams(p, scene) {
    const q = scene || 0;
    const g = this.grids[q];
    const buf = new ArrayBuffer(256);
    const f = new Float32Array(buf);
    const i32 = new Int32Array(buf);
    const u32 = new Uint32Array(buf);
    f.set(g.gridOrigin, 0); f[3] = g.support;
    i32[4] = g.gridDims[0]; i32[5] = g.gridDims[1]; i32[6] = g.gridDims[2];
    u32[7] = g.numCells;
    f.set(p.domainMin, 8); f[11] = p.dt;
    f.set(p.domainMax, 12); f[15] = g.h;
    f.set(p.gravity, 16); f[19] = g.support;
    f.set(p.colliderPos || [0, 0, 0], 20); f[23] = def(p.colliderR, 0);
    f.set(p.colliderVel || [0, 0, 0], 24); f[27] = p.colliderActive ? 1 : 0;
    u32[28] = g.N; f[29] = g.V0; f[30] = p.xsphC; f[31] = p.omega;
    f[32] = p.pbfEps; f[33] = p.sCorrK; f[34] = def(p.time, 0); f[35] = p.contactOmega;
    f[36] = def(p.kDamp, 0.4);
    loadAddon().writeSceneParams(this.handle, q, buf);
  }

  initProfiler() { return true; }

  // sim.js:199-215: one substep with per-launch timing
  profileSubstep(iterations, hasWater) {
    const rows = loadAddon().profileSubstep(this.handle, iterations, !!hasWater);
    return Promise.resolve(rows.map(r => `${r[0]}: ${r[1].toFixed(3)}ms`).join('\n'));
  }

  // sim.js:220: record one substep into the encoder
  encodeSubstep(enc, iterations, hasWater) {
    enc.cmds.push({ sim: this, iterations, hasWater: !!hasWater });
  }

  step(substeps, iterations, hasWater) {
    loadAddon().step(this.handle, substeps, iterations, !!hasWater);
  }

  timeSubsteps(substeps, iterations, hasWater) {
    return loadAddon().timeSubsteps(this.handle, substeps, iterations, !!hasWater);
  }

  // ---- beyond the reference: residual-driven Jacobi, learned warm start,
  // HIP-graph replay and launch shapes (include/xpbi.h)
  // iterations become a cap: stop at residual tol, tested every checkEvery iterations
  setTolerance(tol, checkEvery) { loadAddon().setTolerance(this.handle, tol, checkEvery === undefined ? 4 : checkEvery); }
  // {iterations, residual} of the last substep
  lastIterations() { return loadAddon().lastIterations(this.handle); }
  // Float32Array of the network's 5122 parameters, or null for plain Jacobi
  setWarmstart(params) { loadAddon().setWarmstart(this.handle, params || null); }
  // substeps replayed from a captured HIP graph (same results as step)
  graphStep(substeps, iterations, hasWater) { loadAddon().graphStep(this.handle, substeps, iterations, !!hasWater); }
  setLaunch(blocks) { loadAddon().setLaunch(this.handle, blocks); }
  setNeighbourSlots(slots) { loadAddon().setNeighbourSlots(this.handle, slots); }
  getLaunch() { return loadAddon().getLaunch(this.handle); }
  // times the launch shapes on the live state, keeps the fastest: {blocks, ms}
  autotune(iterations, hasWater, substeps) { return loadAddon().autotune(this.handle, iterations, !!hasWater, substeps || 2); }

  read(field) {
    const ab = loadAddon().read(this.handle, FIELD[field]);
    return (field === 'PERM' || field === 'ID' || field === 'CELLSTART') ? new Uint32Array(ab) : new Float32Array(ab);
  }

  write(field, typed) { loadAddon().write(this.handle, FIELD[field], typed); }

  // main.js:291-312 readback hooks
  readPos() { return Promise.resolve(this.read('POS')); }
  readVel() { return Promise.resolve(this.read('VEL')); }

  destroy() {
    if (this.handle) loadAddon().destroy(this.handle);
    this.handle = null;
  }
}

module.exports = { Simulation, MAT, FIELD, requestDevice, loadAddon, sceneGrid };
