'use strict';
// lifhip.js -- LIFHip, the brain interface of LIFWasm / LIFGpu over the HIP
// engine (src/lifwasm.js:36-87, src/lifgpu.js:185-370 in Lulzx/fly-brain).
//
// Same constructor inputs (graph {indptr, indices, weights, sign}, N, params,
// seed), the same fields (N, p, t, nslots, v, gE, gI, refr, trace, adapt, res,
// bias, thr, drive, spikeCount, drivenSet) and the same methods (step,
// setDrive, setDriveOne, setBias, setThr, addG, pulse, reset, setBackground,
// setParams, flush, destroy, wake, nAwake).  The device holds the state; the
// state arrays here are host copies fetched when read after a step, as
// LIFGpu's shadows are.  drive, thr and bias are written through by the
// setters; direct element writes into those three arrays reach the device on
// flush().  Setter calls queue up and reach the device as one batch
// (fb_apply_deltas) before anything that depends on them, in call order.
//
// Two modes:
//   exact (default)  LIFWasm's semantics: step() runs one step and returns
//                    its fired list; setters act before the next step.
//   live: true       LIFGpu's protocol (src/lifgpu.js:276-346): step() queues
//                    the step and returns the fired list of the last batch;
//                    a batch is submitted after batchSteps steps (8) or
//                    batchMs milliseconds (2), setter calls made during a
//                    batch act after its steps (those made between batches,
//                    before the next one), and the state arrays lag by at
//                    most one batch.  flush() submits.  Batches of steps
//                    run as one engine call, so long enough batches take the
//                    lookahead engine.
// Written in Node-12 syntax (CommonJS, no optional chaining).
const path = require('path');

const native = require(path.join(__dirname, '..', 'lib', 'flybrain_napi.node'));

// DEFAULTS of src/lif.js:6-15
const DEFAULTS = {
  dt: 0.5, vRest: -52, vThresh: -45, vReset: -52, tauM: 20, tauSyn: 5, tRef: 2.2, delay: 1.8,
  wSyn: 0.275, noise: 0, traceTau: 30, adaptInc: 2, adaptTau: 100, depU: 0.2, depTau: 200,
  minSyn: 1, preInh: 0, coba: false, eExc: 0, eInh: -70, inhGain: 1,
};
// fb_params field order (include/flybrain.h)
const PARAM_FIELDS = ['dt', 'vRest', 'vThresh', 'vReset', 'tauM', 'tauSyn', 'tRef', 'delay', 'traceTau',
  'adaptInc', 'adaptTau', 'depU', 'depTau', 'coba', 'eExc', 'eInh', 'bgRate', 'bgAmp', 'preInh'];
const ARRAYS = ['v', 'gE', 'gI', 'refr', 'trace', 'adapt', 'res', 'bias', 'thr', 'drive', 'spikeCount'];
const HOST_OWNED = ['bias', 'thr', 'drive'];
// fb_apply_deltas kinds (include/flybrain.h)
const DELTA = { drive: 0, bias: 1, thr: 2, gE: 3, gI: 4 };
const DELTA_CAP = 65536;   // a full queue is applied at once
const now = () => Number(process.hrtime.bigint()) / 1e6;
const ENGINES = { auto: 0, launch: 1, graph: 2, persistent: 3 };

function paramsVec(p) {
  if (native.paramsLength() !== PARAM_FIELDS.length) throw new Error('LIFHip: fb_params layout mismatch');
  const out = new Float64Array(PARAM_FIELDS.length);
  PARAM_FIELDS.forEach((k, i) => {
    const v = p[k];
    out[i] = k === 'coba' ? (v ? 1 : 0) : (v === undefined || v === null ? 0 : +v);
  });
  return out;
}

function nslotsOf(p) { return Math.max(1, Math.round(p.delay / p.dt)) + 1; }

// src/lif.js:5: unknown, ACh, GABA, Glu, DA, 5HT, OA, histamine
const EXC_SIGN = [0, 1, -1, -1, 1, 1, 1, -1];
const NETWORK_ARRAY_OPTIONS = ['inScale', 'outScale', 'sensoryMask', 'ntSign', 'preSign'];

// LIFNetwork's constructor bake (src/lif.js:17-33): per-edge weights from the
// synapse counts, then inhGain on inhibitory rows; the per-source sign times
// wSyn is what lif.c reads.
function bakeNetwork(N, indptr, indices, counts, nt, p) {
  const SG = p.ntSign || EXC_SIGN, PS = p.preSign, inS = p.inScale, outS = p.outScale;
  const sens = p.sensoryMask || new Uint8Array(N);
  const signOf = (j) => (PS ? PS[j] : SG[nt[j]]);
  const w = new Float32Array(indices.length);
  const sign = new Float32Array(N);
  for (let j = 0; j < N; j++) {
    const og = outS ? outS[j] : 1, s = signOf(j), keepSensory = p.preInh > 0 && s < 0;
    sign[j] = s * p.wSyn;
    for (let k = indptr[j]; k < indptr[j + 1]; k++) {
      const q = indices[k], c = counts[k];
      w[k] = (c >= p.minSyn && (!sens[q] || keepSensory)) ? c * (inS ? inS[q] : 1) * og : 0;
    }
  }
  if (p.inhGain !== 1) {
    for (let j = 0; j < N; j++) if (signOf(j) < 0) for (let k = indptr[j]; k < indptr[j + 1]; k++) w[k] *= p.inhGain;
  }
  return { weights: w, sign: sign };
}

class LIFHip {
  /**
   * @param opts {graph: {indptr, indices, weights, sign, counts?}, N?, params?, seed?, engine?}
   *   engine: 'auto' (default), 'launch', 'graph' or 'persistent'
   */
  constructor(opts) {
    const g = opts.graph;
    const N = opts.N || g.N || (g.indptr.length - 1);
    this.N = N;
    this.p = Object.assign({}, DEFAULTS, opts.params || {});
    this.engine = ENGINES[opts.engine || 'auto'];
    if (this.engine === undefined) throw new Error('LIFHip: unknown engine ' + opts.engine);
    this._graph = native.graphCreate(N, Uint32Array.from(g.indptr), Uint32Array.from(g.indices),
      Float32Array.from(g.weights), Float32Array.from(g.sign), g.counts ? Int32Array.from(g.counts) : null);
    this._h = native.brainCreate(this._graph, paramsVec(this.p), opts.seed === undefined ? 1 : opts.seed);
    this.nslots = nslotsOf(this.p);
    this.t = 0;
    this.drivenSet = new Set();
    this._host = {};
    this._stale = {};
    for (const k of ARRAYS) {
      this._host[k] = k === 'spikeCount' ? new Uint32Array(N) : new Float32Array(N);
      this._stale[k] = true;
    }
    this._fired = new Int32Array(2 * N);
    // setter queue
    this._qIdx = new Int32Array(256);
    this._qKind = new Int32Array(256);
    this._qVal = new Float64Array(256);
    this._qn = 0;
    // LIFGpu's batching protocol
    this.live = !!opts.live;
    this.batchSteps = opts.batchSteps || 8;
    this.batchMs = opts.batchMs === undefined ? 2 : opts.batchMs;
    this._pending = 0;
    this._lastSubmit = now();
    this._lastFired = new Int32Array(0);
  }

  _push(i, kind, v) {
    if (this._qn === this._qIdx.length) {
      const grow = (a, C) => { const b = new C(2 * a.length); b.set(a); return b; };
      this._qIdx = grow(this._qIdx, Int32Array);
      this._qKind = grow(this._qKind, Int32Array);
      this._qVal = grow(this._qVal, Float64Array);
    }
    this._qIdx[this._qn] = i;
    this._qKind[this._qn] = kind;
    this._qVal[this._qn] = v;
    this._qn++;
    if (this._qn >= DELTA_CAP) {
      if (this.live) this._submit(); else this._applyQueue();
    }
  }

  _applyQueue() {
    if (!this._qn) return;
    native.applyDeltas(this._h, this._qIdx, this._qKind, this._qVal, this._qn);
    this._qn = 0;
    for (const k of ['gE', 'gI']) this._stale[k] = true;
  }

  // exact mode: queued setters reach the device; live mode: the open batch is
  // submitted (its steps, then the setters called during it)
  _settle() {
    if (this.live) this._submit(); else this._applyQueue();
  }

  _submit() {
    if (this._pending) {
      const nf = native.step(this._h, this._pending, this.engine);
      this._pending = 0;
      this._deviceMoved();
      const n = native.fired(this._h, this._fired);
      this._lastFired = this._fired.slice(0, Math.min(n, nf));
    }
    this._applyQueue();
    this._lastSubmit = now();
  }

  /** LIFGpu.create's shape (src/lifgpu.js:185): resolves to a ready brain. */
  static create(opts) { return Promise.resolve(new LIFHip(opts)); }

  /**
   * new LIFNetwork(N, indptr, indices, weights, nt, params) (src/lif.js:17-33)
   * on the engine: the same weight and sign bake, lif.c's step.  Gaussian
   * noise is not part of lif.c and is refused.
   */
  static fromNetwork(N, indptr, indices, weights, nt, params, seed, engine) {
    const p = Object.assign({}, DEFAULTS, params || {});
    if (p.noise > 0) throw new Error('LIFHip.fromNetwork: noise > 0 is not part of the lif.c model');
    const baked = bakeNetwork(N, indptr, indices, weights, nt, p);
    const scalar = {};
    for (const k of Object.keys(p)) if (NETWORK_ARRAY_OPTIONS.indexOf(k) < 0) scalar[k] = p[k];
    const b = new LIFHip({ graph: { indptr: indptr, indices: indices, weights: baked.weights, sign: baked.sign },
      N: N, params: scalar, seed: seed, engine: engine });
    b.indptr = indptr; b.indices = indices; b.nt = nt; b.weights = baked.weights;
    b.sensory = p.sensoryMask || new Uint8Array(N);
    return b;
  }

  _array(k) {
    if (this._stale[k]) {
      if (!this.live) this._applyQueue();
      native.getArray(this._h, ARRAYS.indexOf(k), this._host[k]);
      this._stale[k] = false;
    }
    return this._host[k];
  }

  _deviceMoved() { for (const k of ARRAYS) if (HOST_OWNED.indexOf(k) < 0) this._stale[k] = true; }

  get v() { return this._array('v'); }
  get gE() { return this._array('gE'); }
  get gI() { return this._array('gI'); }
  get refr() { return this._array('refr'); }
  get trace() { return this._array('trace'); }
  get adapt() { return this._array('adapt'); }
  get res() { return this._array('res'); }
  get bias() { return this._array('bias'); }
  get thr() { return this._array('thr'); }
  get drive() { return this._array('drive'); }
  get spikeCount() { return this._array('spikeCount'); }
  get nAwake() { return this.N; }
  wake() {}

  /** One step.  Exact mode: runs it and returns its fired list in lif.c
   *  order (a view, valid until the next step).  Live mode: queues it and
   *  returns the fired list of the last submitted batch. */
  step() {
    if (this.live) {
      // setters called while no batch was open act before the new batch
      if (this._pending === 0) this._applyQueue();
      this._pending++;
      this.t += this.p.dt;
      if (this._pending >= this.batchSteps || now() - this._lastSubmit >= this.batchMs) this._submit();
      return this._lastFired;
    }
    this._applyQueue();
    const nf = native.step(this._h, 1, this.engine);
    this.t += this.p.dt;
    this._deviceMoved();
    const n = native.fired(this._h, this._fired);
    return this._fired.subarray(0, Math.min(n, nf));
  }

  /** Many steps off the JS thread; resolves to the spike count of the last step. */
  run(steps) {
    this._settle();
    return native.runAsync(this._h, steps, this.engine).then((nf) => {
      this.t += steps * this.p.dt;
      this._deviceMoved();
      return nf;
    });
  }

  setDriveOne(i, hz) {
    this._push(i, DELTA.drive, hz);
    this._host.drive[i] = hz;
    if (hz > 0) this.drivenSet.add(i); else this.drivenSet.delete(i);
  }

  setDrive(ix, hz) {
    for (const i of ix) this.setDriveOne(i, hz);
  }

  setBias(ix, mv) {
    for (const i of ix) {
      this._push(i, DELTA.bias, mv);
      this._host.bias[i] = mv;
    }
  }

  setThr(i, mv) {
    this._push(i, DELTA.thr, mv);
    this._host.thr[i] = mv;
  }

  addG(i, e, ii) {
    this._push(i, DELTA.gE, e);
    this._push(i, DELTA.gI, ii || 0);
  }

  pulse(ix, mv) {
    for (const i of ix) this._push(i, DELTA.gE, mv);
  }

  reset() {
    this._settle();
    native.reset(this._h);
    this._deviceMoved();
  }

  setBackground(hz, mv) {
    this._settle();
    native.setBackground(this._h, hz, mv);
  }

  setParams(q) {
    this._settle();
    Object.assign(this.p, q);
    native.setParams(this._h, paramsVec(this.p));
    this.nslots = nslotsOf(this.p);
  }

  /** Pushes direct element writes of drive / thr / bias to the device and
   *  drops the cached state arrays. */
  flush() {
    this._settle();
    for (const k of HOST_OWNED) if (!this._stale[k]) native.setArray(this._h, ARRAYS.indexOf(k), this._host[k]);
    this._deviceMoved();
  }

  /** Work counters: edge events, synapse events, spikes, steps. */
  counters() {
    if (!this.live) this._applyQueue();
    const c = native.counters(this._h);
    return { edgeEvents: c[0], synapseEvents: c[1], spikes: c[2], steps: c[3] };
  }

  get rng() { return native.scalars(this._h)[0]; }

  destroy() {
    this._qn = 0;
    this._pending = 0;
    if (this._h) { native.brainDestroy(this._h); this._h = null; }
    if (this._graph) { native.graphRelease(this._graph); this._graph = null; }
  }
}

module.exports = { LIFHip, DEFAULTS, PARAM_FIELDS, EXC_SIGN, bakeNetwork, native };
