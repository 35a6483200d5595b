'use strict';
// Runs one brain through the N-API addon and the LIFHip class, for the parity
// test: reads a JSON case (graph, params, seed, ops) and writes per-step spike
// counts, the raster, spike counts and the RNG state as JSON.
const fs = require('fs');
const path = require('path');
const { LIFHip, native } = require(path.join(__dirname, '..', '..', 'fly-brain_amd', 'js', 'lifhip.js'));

const spec = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
if (spec.exportsOnly) {
  process.stdout.write(JSON.stringify({ exports: Object.keys(native).sort(), paramsLength: native.paramsLength() }));
  process.exit(0);
}
const brain = new LIFHip({ graph: spec.graph, N: spec.N, params: spec.params, seed: spec.seed, engine: spec.engine,
  live: spec.live, batchSteps: spec.batchSteps, batchMs: spec.batchMs });
const nfired = [];
const raster = [];
const run = async () => {
  for (const op of spec.ops) {
    if (op[0] === 'setDriveOne') brain.setDriveOne(op[1], op[2]);
    else if (op[0] === 'setDrive') brain.setDrive(op[1], op[2]);
    else if (op[0] === 'setBias') brain.setBias(op[1], op[2]);
    else if (op[0] === 'setThr') brain.setThr(op[1], op[2]);
    else if (op[0] === 'pulse') brain.pulse(op[1], op[2]);
    else if (op[0] === 'addG') brain.addG(op[1], op[2], op[3]);
    else if (op[0] === 'setThrMany') { for (const [i, mv] of op[1]) brain.setThr(i, mv); }
    else if (op[0] === 'flush') brain.flush();
    else if (op[0] === 'steps') {
      for (let s = 0; s < op[1]; s++) {
        const f = brain.step();
        nfired.push(f.length);
        for (let k = 0; k < f.length; k++) raster.push(f[k]);
      }
    } else if (op[0] === 'run') {
      await brain.run(op[1]);
    } else throw new Error('unknown op ' + op[0]);
  }
  const out = {
    nfired, raster, t: brain.t, rng: brain.rng,
    spikeCount: Array.from(brain.spikeCount), v: Array.from(brain.v),
    drivenSet: Array.from(brain.drivenSet), counters: brain.counters(),
  };
  brain.destroy();
  process.stdout.write(JSON.stringify(out));
};
run().catch((e) => { console.error(e.stack || String(e)); process.exit(1); });
