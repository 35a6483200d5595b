'use strict';
// Builds LIFHip.fromNetwork from a JSON spec {N, indptr, indices, weights, nt,
// params, seed, engine, driven, rate, steps} and prints the per-step spike
// counts and the final spike counts as JSON.
const fs = require('fs');
const path = require('path');
const { LIFHip } = require(path.join(__dirname, '..', '..', 'fly-brain_amd', 'js', 'lifhip.js'));

const s = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
const b = LIFHip.fromNetwork(s.N, Uint32Array.from(s.indptr), Uint32Array.from(s.indices),
  Float32Array.from(s.weights), Uint8Array.from(s.nt), s.params, s.seed, s.engine);
b.setDrive(s.driven, s.rate);
const nf = [];
for (let t = 0; t < s.steps; t++) nf.push(b.step().length);
const sc = Array.from(b.spikeCount);
b.destroy();
process.stdout.write(JSON.stringify({ nfired: nf, spikeCount: sc }));
