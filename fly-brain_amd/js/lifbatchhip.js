'use strict';
// lifbatchhip.js -- LIFBatchHip, B ensemble members of one brain model over
// one graph on the HIP engine: the one-brain-per-worker ensembles of
// src/brainsetup.js:8 and scripts/lif_ensemble.mjs:16-35 (Lulzx/fly-brain),
// and the lane ensemble of kernel.bend:293-378.
//
// Member b is the brain LIFHip would make with seed seeds[b] and weights
// f32(count * gains[b][type(pre)]) (writeGraph's outScale rule,
// src/lifwasm.js:23,31, with a per-type gain); members share the driven set
// and rates, and each has its own thresholds and biases (silencing is
// setThr(i, 1e6), scripts/lif_ensemble.mjs:35) and run length.
// Written in Node-12 syntax (CommonJS, no optional chaining).
const { native, DEFAULTS, PARAM_FIELDS } = require('./lifhip.js');

const ARRAYS = ['v', 'gE', 'gI', 'refr', 'trace', 'adapt', 'res', 'bias', 'thr', 'drive', 'spikeCount'];
const ENGINES = { auto: 0, launch: 1, graph: 2, persistent: 3 };
const MEMBER = { thr: 8, bias: 7 };     // FB_THR, FB_BIAS

function paramsVec(p) {
  const out = new Float64Array(PARAM_FIELDS.length);
  PARAM_FIELDS.forEach((k, i) => {
    const v = p[k];
    out[i] = k === 'coba' ? (v ? 1 : 0) : (v === undefined || v === null ? 0 : +v);
  });
  return out;
}

class LIFBatchHip {
  /**
   * @param opts {graph: {indptr, indices, weights, sign, counts}, N?, params?,
   *   seeds (B of them), typeOf[N], gains[B][nTypes] (rows) or a flat
   *   Float32Array of B * nTypes, engine?}
   */
  constructor(opts) {
    const g = opts.graph;
    const N = opts.N || (g.indptr.length - 1);
    if (!g.counts) throw new Error('LIFBatchHip: the graph needs its synapse counts');
    this.N = N;
    this.p = Object.assign({}, DEFAULTS, opts.params || {});
    this.B = opts.seeds.length;
    const flat = opts.gains instanceof Float32Array ? opts.gains : Float32Array.from([].concat.apply([], opts.gains.map((r) => Array.from(r))));
    this.nTypes = flat.length / this.B;
    this._graph = native.graphCreate(N, Uint32Array.from(g.indptr), Uint32Array.from(g.indices),
      Float32Array.from(g.weights), Float32Array.from(g.sign), Int32Array.from(g.counts));
    this._h = native.batchCreate(this._graph, paramsVec(this.p), Float64Array.from(opts.seeds),
      Int32Array.from(opts.typeOf), this.nTypes, flat);
    if (opts.engine) native.batchSetEngine(this._h, ENGINES[opts.engine]);
    this.drivenSet = new Set();
  }

  setDrive(ix, hz) {
    for (let k = 0; k < ix.length; k++) { if (hz > 0) this.drivenSet.add(ix[k]); else this.drivenSet.delete(ix[k]); }
    native.batchSetDrive(this._h, Int32Array.from(ix), hz);
  }
  setDriveOne(i, hz) { this.setDrive([i], hz); }
  /** member b's threshold offsets on a list of neurons */
  setThr(b, ix, mv) { native.batchSetMember(this._h, MEMBER.thr, b, Int32Array.from(ix), mv); }
  setBias(b, ix, mv) { native.batchSetMember(this._h, MEMBER.bias, b, Int32Array.from(ix), mv); }
  /** silences member b's neurons (setThr(i, 1e6), scripts/lif_ensemble.mjs:35) */
  silence(b, ix) { this.setThr(b, ix, 1e6); }
  /** steps[b] more steps for member b (negative: no limit) */
  setSteps(steps) { native.batchSetSteps(this._h, Int32Array.from(steps)); }
  run(nsteps) { native.batchRun(this._h, nsteps); }

  /** member b's state array */
  array(name, b) {
    const out = name === 'spikeCount' ? new Uint32Array(this.N) : new Float32Array(this.N);
    native.batchGetArray(this._h, ARRAYS.indexOf(name), b, name === 'spikeCount' ? new Float32Array(out.buffer) : out);
    return out;
  }
  spikeCount(b) { return this.array('spikeCount', b); }

  /** lane mode (cordens.bend): the type gain scales the sign; laneDrive[B][N] per-member rates */
  setLanes(signGain, laneDrive) {
    let flat = null;
    if (laneDrive) flat = laneDrive instanceof Float32Array ? laneDrive : Float32Array.from([].concat.apply([], laneDrive.map((r) => Array.from(r))));
    native.batchSetLanes(this._h, signGain ? 1 : 0, flat);
  }
  setRng(states) { native.batchSetRng(this._h, Uint32Array.from(states)); }
  get rng() { const out = new Uint32Array(this.B); native.batchRng(this._h, out); return out; }
  counters() {
    const c = native.batchCounters(this._h);
    return { edgeEvents: c[0], synapseEvents: c[1], spikes: c[2], steps: c[3] };
  }
  stat(name) { return native.batchStat(this._h, name); }
  destroy() {
    if (this._h) { native.batchDestroy(this._h); this._h = null; }
    if (this._graph) { native.graphRelease(this._graph); this._graph = null; }
  }
}

module.exports = { LIFBatchHip };
