'use strict';
// Checks LIFHip's LIFNetwork bake (fly-brain_amd/js/lifhip.js: bakeNetwork)
// against the baked weights of the golden LIFNetwork runs.  Prints JSON.
const fs = require('fs');
const path = require('path');
const { bakeNetwork, DEFAULTS } = require(path.join(__dirname, '..', '..', 'fly-brain_amd', 'js', 'lifhip.js'));

const gold = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
const out = {};
for (const name of Object.keys(gold)) {
  const c = gold[name], g = c.graph, ex = c.extra || {};
  const p = Object.assign({}, DEFAULTS, c.params);
  if (ex.inScale) p.inScale = Float32Array.from(ex.inScale);
  if (ex.outScale) p.outScale = Float32Array.from(ex.outScale);
  if (ex.sensoryMask) p.sensoryMask = Uint8Array.from(ex.sensoryMask);
  const b = bakeNetwork(g.N, Uint32Array.from(g.indptr), Uint32Array.from(g.indices), Float32Array.from(g.weights),
    Uint8Array.from(g.nt), p);
  const bits = new Uint32Array(b.weights.buffer);
  let same = bits.length === c.weights_baked.length;
  for (let k = 0; same && k < bits.length; k++) same = bits[k] === c.weights_baked[k];
  out[name] = same;
}
process.stdout.write(JSON.stringify(out));
