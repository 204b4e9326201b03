'use strict';
// Command-line driver of the JavaScript host.
//   node run.js hash  <scene name | c2 | headline> <seed>   sha256 of the generated scene arrays (no GPU)
//   node run.js frames <scene name> <seed> <frames> <out.bin>  run frames on HIP device 0 and write
//                                                         POS (float32) followed by ID (uint32)
//   node run.js exports                                   list the addon's functions (no GPU call)
const crypto = require('crypto');
const fs = require('fs');
const scenes = require('./scenes');

function makeScene(name, seed) {
  if (name === 'c2') return scenes.configC2(seed);
  if (name === 'headline') return scenes.configHeadline(seed);
  const f = scenes.SCENES[name];
  if (!f) throw new Error('unknown scene ' + name);
  return f(seed);
}

function hashScene(s) {
  const h = crypto.createHash('sha256');
  h.update(Buffer.from(s.positions.buffer));
  h.update(Buffer.from(s.tags.buffer));
  h.update(Buffer.from(s.states.buffer));
  if (s.velocities) h.update(Buffer.from(s.velocities.buffer));
  return h.digest('hex');
}

const [cmd, name, seedS, framesS, out] = process.argv.slice(2);
if (cmd === 'hash') {
  const s = makeScene(name, parseInt(seedS, 10));
  console.log(JSON.stringify({ N: s.positions.length / 3, sha256: hashScene(s) }));
} else if (cmd === 'exports') {
  const { loadAddon } = require('./sim');
  console.log(JSON.stringify(Object.keys(loadAddon()).sort()));
} else if (cmd === 'frames') {
  const { Simulation, requestDevice } = require('./sim');
  const { FrameDriver } = require('./frame');
  const s = makeScene(name, parseInt(seedS, 10));
  const device = requestDevice({ ordinal: 0 });
  const sim = new Simulation(device, s);
  const drv = new FrameDriver(device, sim, s);
  const frames = parseInt(framesS, 10);
  for (let f = 0; f < frames; f++) drv.frame(1 / 60);
  const pos = sim.read('POS');
  const id = sim.read('ID');
  fs.writeFileSync(out, Buffer.concat([Buffer.from(pos.buffer), Buffer.from(id.buffer)]));
  console.log(JSON.stringify({ N: sim.N, frames, simTime: drv.simTime }));
  sim.destroy();
} else {
  console.error('usage: node run.js hash|frames|exports ...');
  process.exit(2);
}
