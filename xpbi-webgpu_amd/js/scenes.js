'use strict';
// Scene definitions (the reference's web/js/scenes.js:7-176) with a seeded,
// counter-based PRNG in place of Math.random: lowbias32 of a counter, computed
// with Math.imul, bit-identical to xpbi-webgpu_amd/xpbi/scenes.py.  Positions
// are formed in double in the reference's operation order and rounded to f32
// on upload.

const { MAT } = require('./sim');

function lowbias32(x) {
  x = x >>> 0;
  x ^= x >>> 16;
  x = Math.imul(x, 0x7feb352d) >>> 0;
  x ^= x >>> 15;
  x = Math.imul(x, 0x846ca68b) >>> 0;
  x ^= x >>> 16;
  return x >>> 0;
}
function streamKey(seed, stream) {
  const k = lowbias32(seed >>> 0);
  return lowbias32((k + Math.imul(stream, 0x632be5ab)) >>> 0);
}
function randU32(seed, stream, counter) {
  return lowbias32((Math.imul(counter >>> 0, 0x9e3779b9) + streamKey(seed, stream)) >>> 0);
}
function randUnit(seed, stream, counter) {
  return (randU32(seed, stream, counter) >>> 8) * (1 / 16777216);
}
class Rng {
  constructor(seed, stream) { this.seed = seed; this.stream = stream; this.k = 0; this.key = streamKey(seed, stream); }
  next() {
    const u = lowbias32((Math.imul(this.k >>> 0, 0x9e3779b9) + this.key) >>> 0);
    this.k++;
    return (u >>> 8) * (1 / 16777216);
  }
}

// scenes.js:7-20
function sampleBox(out, min, max, h, matIdx, rng, jitter, state0) {
  jitter = jitter === undefined ? 0.2 : jitter;
  state0 = state0 || 0;
  const n = [0, 1, 2].map(i => Math.max(1, Math.round((max[i] - min[i]) / h)));
  let cnt = 0;
  for (let k = 0; k < n[2]; k++)
    for (let j = 0; j < n[1]; j++)
      for (let i = 0; i < n[0]; i++) {
        const ux = rng.next(), uy = rng.next(), uz = rng.next();
        out.pos.push(
          min[0] + (i + 0.5) * h + (ux - 0.5) * jitter * h,
          min[1] + (j + 0.5) * h + (uy - 0.5) * jitter * h,
          min[2] + (k + 0.5) * h + (uz - 0.5) * jitter * h);
        out.mat.push(matIdx);
        out.state.push(state0);
        cnt++;
      }
  return cnt;
}

// scenes.js:22-38 (distance as sqrt of the summed squares)
function sampleSphere(out, c, r, h, matIdx, rng, jitter, state0) {
  jitter = jitter === undefined ? 0.2 : jitter;
  state0 = state0 || 0;
  const n = Math.ceil(2 * r / h);
  let cnt = 0;
  for (let k = 0; k <= n; k++)
    for (let j = 0; j <= n; j++)
      for (let i = 0; i <= n; i++) {
        const p = [c[0] - r + i * h, c[1] - r + j * h, c[2] - r + k * h];
        const dx = p[0] - c[0], dy = p[1] - c[1], dz = p[2] - c[2];
        const d = Math.sqrt(dx * dx + dy * dy + dz * dz);
        if (d <= r) {
          const ux = rng.next(), uy = rng.next(), uz = rng.next();
          out.pos.push(p[0] + (ux - 0.5) * jitter * h, p[1] + (uy - 0.5) * jitter * h, p[2] + (uz - 0.5) * jitter * h);
          out.mat.push(matIdx);
          out.state.push(state0);
          cnt++;
        }
      }
  return cnt;
}

// scenes.js:40-80
function dpAlpha(phiDeg) {
  const s = Math.sin(phiDeg * Math.PI / 180);
  return Math.sqrt(2 / 3) * 2 * s / (3 - s);
}
const SAND = (phi, cohesion, E, nu, density) => ({
  kind: MAT.DP, E: E === undefined ? 2000 : E, nu: nu === undefined ? 0.3 : nu,
  density: density === undefined ? 1.0 : density, p0: dpAlpha(phi === undefined ? 35 : phi), p1: cohesion || 0,
  color: [0.85, 0.63, 0.28], boundFriction: 0.5,
});
const SNOW_NACC = () => ({ kind: MAT.NACC, E: 2000, nu: 0.3, density: 4.0, p0: -0.008, p1: 0.05, p2: 5, p3: 1.85,
  color: [0.93, 0.95, 0.98], boundFriction: 0.35 });
const PLASTICINE = (sigY, E, nu, density) => ({ kind: MAT.VM, E: E === undefined ? 1200 : E,
  nu: nu === undefined ? 0.42 : nu, density: density === undefined ? 2.0 : density, p0: sigY === undefined ? 30 : sigY,
  color: [0.36, 0.76, 0.72], boundFriction: 0.4 });
const GOO = hExp => ({ kind: MAT.HB, E: 2250, nu: 0.125, density: 100.0, p0: 45, p1: hExp === undefined ? 1.0 : hExp,
  p2: 25, color: [0.91, 0.45, 0.62], boundFriction: 0.2 });
const SNOW_CLASSIC = () => ({ kind: MAT.SNOW, E: 9000, nu: 0.2, density: 4.0, p0: 3.5e-2, p1: 8e-3, p2: 10,
  color: [0.96, 0.97, 1.0], boundFriction: 0.35 });
const WATER = density => ({ kind: MAT.WATER, E: 100, nu: 0.3, density: density === undefined ? 1.0 : density,
  color: [0.25, 0.55, 0.95], boundFriction: 0.0 });
const ELASTIC = () => ({ kind: MAT.ELASTIC, E: 900, nu: 0.3, density: 2.0, color: [0.62, 0.52, 0.86], boundFriction: 0.3 });

// scenes.js:82-92
function base(h) {
  return {
    h, domainMin: [0, 0, 0], domainMax: [1, 1, 1], gravity: [0, -9.8, 0],
    substeps: 8, iterations: 3, timeScale: 0.42,
    omega: 0.5, contactOmega: 0.4, xsphC: 0.01, pbfEps: 100.0, sCorrK: 0.0001, kDamp: 0.4,
  };
}

function newOut() { return { pos: [], mat: [], state: [] }; }

function finish(s, out, materials, seed, hasWater, velocities, info) {
  const n = out.mat.length;
  const tags = new Float32Array(n);
  for (let i = 0; i < n; i++) tags[i] = out.mat[i] + randUnit(seed, 99, i) * 0.49; // sim.js:63, seeded
  const r = Object.assign({}, s, {
    positions: Float32Array.from(out.pos), matIds: Uint32Array.from(out.mat), tags,
    states: Float32Array.from(out.state), materials, hasWater, info, seed,
  });
  if (velocities) r.velocities = Float32Array.from(velocities);
  return r;
}

const SCENES = {
  'Sand: column collapse': (seed = 0) => {
    const s = base(0.016);
    const out = newOut(), rng = new Rng(seed, 0);
    sampleBox(out, [0.36, 0.015, 0.36], [0.64, 0.62, 0.64], s.h, 0, rng);
    return finish(s, out, [SAND(35)], seed, false, null, 'Sand: column collapse');
  },
  'Sand: two blocks collide': (seed = 0) => {
    const s = base(0.017);
    const out = newOut(), rng = new Rng(seed, 0);
    const n1 = sampleBox(out, [0.08, 0.015, 0.35], [0.38, 0.42, 0.65], s.h, 0, rng);
    sampleBox(out, [0.62, 0.015, 0.35], [0.92, 0.42, 0.65], s.h, 0, rng);
    const vel = new Float64Array(out.mat.length * 3);
    for (let i = 0; i < out.mat.length; i++) vel[i * 3] = i < n1 ? 1.6 : -1.6;
    return finish(s, out, [SAND(35)], seed, false, vel, 'Sand: two blocks collide');
  },
  'Snow: snowball smash': (seed = 0) => {
    const s = base(0.018);
    s.substeps = 9; s.iterations = 3; s.timeScale = 0.45;
    const m = SNOW_NACC();
    const out = newOut(), rng = new Rng(seed, 0);
    const n1 = sampleBox(out, [0.1, 0.015, 0.1], [0.9, 0.14, 0.9], s.h, 0, rng, 0.2, m.p0);
    sampleSphere(out, [0.2, 0.55, 0.5], 0.11, s.h, 0, rng, 0.2, m.p0);
    const vel = new Float64Array(out.mat.length * 3);
    for (let i = n1; i < out.mat.length; i++) { vel[i * 3] = 2.6; vel[i * 3 + 1] = -3.2; }
    return finish(s, out, [m], seed, false, vel, 'Snow: snowball smash');
  },
  'Plasticine: drop & poke': (seed = 0) => {
    const s = base(0.013);
    s.substeps = 8; s.iterations = 6; s.timeScale = 0.4;
    const out = newOut(), rng = new Rng(seed, 0);
    sampleSphere(out, [0.5, 0.4, 0.5], 0.16, s.h, 0, rng);
    return finish(s, out, [PLASTICINE(80)], seed, false, null, 'Plasticine: drop & poke');
  },
  'Goo: Herschel-Bulkley': (seed = 0) => {
    const s = base(0.013);
    s.substeps = 9; s.iterations = 3; s.timeScale = 0.5;
    const out = newOut(), rng = new Rng(seed, 0);
    sampleSphere(out, [0.5, 0.42, 0.5], 0.17, s.h, 0, rng);
    return finish(s, out, [GOO(1.0)], seed, false, null, 'Goo: Herschel-Bulkley');
  },
  'Snow: classic (Stomakhin)': (seed = 0) => {
    const s = base(0.013);
    s.substeps = 10; s.timeScale = 0.38; s.iterations = 4;
    const out = newOut(), rng = new Rng(seed, 0);
    sampleSphere(out, [0.35, 0.45, 0.42], 0.1, s.h, 0, rng, 0.2, 1.0);
    sampleSphere(out, [0.65, 0.5, 0.58], 0.1, s.h, 0, rng, 0.2, 1.0);
    const vel = new Float64Array(out.mat.length * 3);
    for (let i = 0; i < out.mat.length; i++) vel[i * 3] = out.pos[i * 3] < 0.5 ? 1.5 : -1.5;
    return finish(s, out, [SNOW_CLASSIC()], seed, false, vel, 'Snow: classic (Stomakhin)');
  },
  'Dam breach: water + sand': (seed = 0) => {
    const s = base(0.019);
    s.substeps = 6; s.iterations = 3; s.timeScale = 0.45;
    const out = newOut(), rng = new Rng(seed, 0);
    sampleBox(out, [0.42, 0.015, 0.05], [0.6, 0.34, 0.95], s.h, 0, rng);
    sampleBox(out, [0.03, 0.015, 0.05], [0.36, 0.5, 0.95], s.h, 1, rng);
    return finish(s, out, [SAND(30, 0.004), WATER()], seed, true, null, 'Dam breach: water + sand');
  },
  'Elastic: StVK block': (seed = 0) => {
    const s = base(0.014);
    s.substeps = 10; s.iterations = 4; s.timeScale = 0.5;
    const out = newOut(), rng = new Rng(seed, 0);
    sampleBox(out, [0.38, 0.3, 0.38], [0.62, 0.54, 0.62], s.h, 0, rng);
    return finish(s, out, [ELASTIC()], seed, false, null, 'Elastic: StVK block');
  },
};

// benchmark workloads (seeded-jitter DP columns)
function sandColumn(nx, ny, nz, r, domain, seed, phi, dt, iterations, jitter) {
  phi = phi === undefined ? 30 : phi;
  const s = base(r);
  s.gravity = [0, -9.81, 0];
  s.dt = dt === undefined ? 1 / 6000 : dt;
  s.iterations = iterations === undefined ? 20 : iterations;
  s.domainMin = [0, 0, 0];
  s.domainMax = domain.slice();
  const ext = [nx * r, ny * r, nz * r];
  const mn = [0.5 * (domain[0] - ext[0]), 0.0, 0.5 * (domain[2] - ext[2])];
  const mx = [mn[0] + ext[0], ext[1], mn[2] + ext[2]];
  const out = newOut(), rng = new Rng(seed, 0);
  sampleBox(out, mn, mx, r, 0, rng, jitter === undefined ? 0.1 : jitter);
  return finish(s, out, [SAND(phi, 0, 3.5e5, 0.3, 2000)], seed, false, null, `DP sand column ${nx}x${ny}x${nz}`);
}
const sandColumn128k = (seed = 2) => sandColumn(40, 80, 40, 0.01, [2, 1, 2], seed);
const sandColumn1m = (seed = 1) => sandColumn(80, 160, 80, 0.005, [1.5, 1, 1.5], seed);

module.exports = {
  SCENES, sampleBox, sampleSphere, base, dpAlpha, SAND, SNOW_NACC, PLASTICINE, GOO, SNOW_CLASSIC, WATER, ELASTIC,
  sandColumn, sandColumn128k, sandColumn1m, lowbias32, randU32, randUnit, Rng,
};
