// Writes tests/golden/lifnet_golden.json from the reference's own LIFNetwork
// (src/lif.js), with Math.random seeded by mulberry32 the way the reference's
// ensemble scripts seed it (scripts/lif_ensemble.mjs:8-12).
//
// Run where /root/reference exists:  node tests/golden/make_lifnet_golden.mjs
// The fixture holds the inputs (a small graph and the calls made) and the
// outputs (every step's fired list and the final state as f32 bit patterns).
import { LIFNetwork } from '/root/reference/src/lif.js';
import { writeFileSync } from 'fs';

function mulberry32(seed) {
  let rs = seed;
  return () => { rs |= 0; rs = rs + 0x6D2B79F5 | 0; let t = Math.imul(rs ^ rs >>> 15, 1 | rs); t = t + Math.imul(t ^ t >>> 7, 61 | t) ^ t; return ((t ^ t >>> 14) >>> 0) / 4294967296; };
}

// a small graph from a 32-bit LCG (numbers are plain data in the fixture)
function graph(N, deg, seed) {
  let s = seed >>> 0;
  const next = () => { s = (Math.imul(s, 1664525) + 1013904223) >>> 0; return s; };
  const indptr = [0], indices = [], weights = [], nt = [];
  for (let j = 0; j < N; j++) {
    const d = 1 + next() % (2 * deg);
    const row = new Set();
    while (row.size < d) { const q = next() % N; if (q !== j) row.add(q); }
    for (const q of [...row].sort((a, b) => a - b)) { indices.push(q); weights.push(1 + next() % 40); }
    indptr.push(indices.length);
    nt.push(next() % 8);
  }
  return { N, indptr, indices, weights, nt };
}

const f32bits = (a) => Array.from(new Uint32Array(Float32Array.from(a).buffer));

const CASES = {
  cuba_default: { N: 300, deg: 12, gseed: 11, rseed: 101, steps: 400, params: {},
    ops: [['setDrive', 0, 30, 150.0]] },
  cuba_recurrent: { N: 300, deg: 12, gseed: 11, rseed: 111, steps: 400, params: { wSyn: 1.0 },
    ops: [['setDrive', 0, 30, 150.0]] },
  coba_dep_adapt: { N: 300, deg: 14, gseed: 12, rseed: 202, steps: 400,
    params: { dt: 0.1, coba: true, eInh: -80, wSyn: 2.0, depU: 0.3, adaptInc: 1.5, inhGain: 1.7, minSyn: 2 },
    ops: [['setDrive', 10, 40, 120.0], ['setBias', 100, 110, 3.0]] },
  preinh_scaled: { N: 260, deg: 10, gseed: 13, rseed: 303, steps: 300,
    params: { dt: 0.2, preInh: 0.5, wSyn: 2.5 }, scaled: true,
    ops: [['setDrive', 0, 40, 200.0], ['pulse', 50, 60, 8.0]] },
  noise: { N: 200, deg: 10, gseed: 14, rseed: 404, steps: 200, params: { noise: 2.0 },
    ops: [['setDrive', 0, 20, 100.0]] },
  setters_mid_run: { N: 240, deg: 12, gseed: 15, rseed: 505, steps: 300, params: { wSyn: 3.0 },
    ops: [['setDrive', 0, 24, 180.0]],
    mid: { at: 150, ops: [['setThr', 5, 1e6], ['addG', 30, 4.0, -1.0], ['setDriveOne', 3, 0.0], ['reset']] } },
};

const out = {};
for (const [name, c] of Object.entries(CASES)) {
  const g = graph(c.N, c.deg, c.gseed);
  const params = { ...c.params };
  const extra = {};
  if (c.scaled) {
    let s = 7;
    const next = () => { s = (Math.imul(s, 1664525) + 1013904223) >>> 0; return s; };
    extra.inScale = Array.from({ length: c.N }, () => 0.5 + (next() % 1000) / 1000);
    extra.outScale = Array.from({ length: c.N }, () => 0.5 + (next() % 1000) / 1000);
    extra.sensoryMask = Array.from({ length: c.N }, (_, i) => (i < 40 ? 1 : 0));
    params.inScale = Float32Array.from(extra.inScale);
    params.outScale = Float32Array.from(extra.outScale);
    params.sensoryMask = Uint8Array.from(extra.sensoryMask);
  }
  Math.random = mulberry32(c.rseed);
  const net = new LIFNetwork(g.N, Uint32Array.from(g.indptr), Uint32Array.from(g.indices),
                             Float32Array.from(g.weights), Uint8Array.from(g.nt), params);
  const apply = (op) => {
    const [f, ...a] = op;
    if (f === 'setDrive' || f === 'setBias' || f === 'pulse') {
      const ix = []; for (let i = a[0]; i < a[1]; i++) ix.push(i);
      net[f](ix, a[2]);
    } else if (f === 'reset') net.reset();
    else net[f](...a);
  };
  c.ops.forEach(apply);
  const fired = [];
  for (let t = 0; t < c.steps; t++) {
    if (c.mid && t === c.mid.at) c.mid.ops.forEach(apply);
    fired.push(Array.from(net.step()));
  }
  out[name] = {
    graph: g, params: c.params, extra, rseed: c.rseed, steps: c.steps, ops: c.ops, mid: c.mid || null,
    weights_baked: f32bits(net.weights),
    fired,
    final: { v: f32bits(net.v), gE: f32bits(net.gE), gI: f32bits(net.gI), refr: f32bits(net.refr),
             trace: f32bits(net.trace), adapt: f32bits(net.adapt), res: f32bits(net.res),
             spikeCount: Array.from(net.spikeCount), nAwake: net.nAwake, t: net.t },
  };
}
writeFileSync(new URL('./lifnet_golden.json', import.meta.url), JSON.stringify(out));
console.log('wrote', Object.keys(out).length, 'cases');
