'use strict';
// lifdiffhip.js -- LIFDiffHip, the differentiable brain of LIFDiff over the
// HIP engine (src/lifdiff.js:84-686 in Lulzx/fly-brain), and LIFDiffBatchHip,
// several of them stepped in lockstep over one graph.
//
// Same constructor inputs (data {N, indptr, indices, weights, nt}, opts with
// LIFDiff's options and arrays: sizeLog, thrMask, biasMask, sensoryMask,
// inScale, logGain, outScale, thrOffset, edgeGain, preSign, driveIdx,
// driveEpoch, surrogateBeta as a number or a Float32Array, soft, driveSoft,
// exact32, adjClip), the same methods (setDrive, setDriveOne, reset,
// forward(steps, {seed, record, onEpoch, spikeTrain}), backward(tape, dLdSpike, {truncate,
// lossFrom, dLdSpikePerEpoch})) and results ({params, logGain, drive,
// truncate, clamped}).  The device holds the state and the tape; the tape
// object returned by forward() names the run it belongs to, and only the last
// forward pass of a network can be differentiated (the reference keeps every
// tape it returned; here a second forward replaces the first).
//
// onEpoch: the reference calls it at every epoch start of the forward pass, and
// it may call setDrive / setDriveOne.  Here it is called for every epoch before
// the pass, and the rates of driveIdx it leaves are the pass's drive schedule:
// the same run whenever the callback does not read the network's own state
// (the visual chain's coupling does not).  spikeTrain[t] (an index list or
// null per step) forces those neurons at step t of that pass.
// Written in Node-12 syntax (CommonJS, no optional chaining).
const { native, EXC_SIGN } = require('./lifhip.js');

// LIFDiff's options and defaults (src/lifdiff.js:88-109); `checkpoint` is the
// replay segment length (the reference fixes 16)
const DIFF_DEFAULTS = {
  dt: 0.5, vRest: -52, vThresh: -45, vReset: -52, tauM: 20, tauSyn: 5, tRef: 2.2, delay: 1.8,
  wSyn: 0.3, adaptInc: 0, adaptTau: 100, depU: 0, depTau: 200, minSyn: 1,
  coba: true, eExc: 0, eInh: -70, inhGain: 1, sizeAlpha: 0, maxSizeScale: 20, boostCap: 1,
  kcThreshold: 0, laminaBias: 0, surrogateBeta: 2.0, adjClip: 0, soft: false, driveSoft: false,
  exact32: false, checkpoint: 16,
};
// fb_diff_params field order (include/flybrain.h): doubles, then soft, checkpoint, exact32
const DIFF_FIELDS = ['dt', 'vRest', 'vThresh', 'vReset', 'tauM', 'tauSyn', 'tRef', 'delay', 'wSyn', 'adaptInc',
  'adaptTau', 'depU', 'depTau', 'eExc', 'eInh', 'inhGain', 'kcThreshold', 'laminaBias', 'surrogateBeta', 'adjClip',
  'sizeAlpha', 'maxSizeScale', 'boostCap'];
const DIFF_PARAMS = ['wSyn', 'sizeAlpha', 'inhGain', 'eInh', 'vThresh', 'kcThreshold', 'laminaBias', 'adaptInc', 'depU'];
const INPUT = { outScale: 0, logGain: 1, thrOffset: 2, sizeLog: 3, drive: 4, inScale: 5, surrogateBeta: 6 };
const STATE = { v: 0, gE: 1, gI: 2, adapt: 3, res: 4, refr: 5, spikeCount: 6, inScale: 7 };
const FORCED = 0x80000000;

function diffParamsVec(p) {
  if (native.diffParamsLength() !== DIFF_FIELDS.length + 3) throw new Error('LIFDiffHip: fb_diff_params layout mismatch');
  const out = new Float64Array(DIFF_FIELDS.length + 3);
  DIFF_FIELDS.forEach((k, i) => {
    const v = p[k];
    out[i] = (k === 'surrogateBeta' && typeof v !== 'number') ? 2.0 : +v;
  });
  out[DIFF_FIELDS.length] = p.soft ? 1 : 0;
  out[DIFF_FIELDS.length + 1] = p.checkpoint;
  out[DIFF_FIELDS.length + 2] = p.exact32 ? 1 : 0;
  return out;
}

class LIFDiffHip {
  /**
   * @param data {N, indptr, indices, weights (synapse counts), nt}
   * @param opts LIFDiff's options (see the file header)
   */
  constructor(data, opts) {
    opts = opts || {};
    const N = data.N;
    const p = this.p = Object.assign({}, DIFF_DEFAULTS, opts);
    this.N = N;
    this.nslots = Math.max(1, Math.round(p.delay / p.dt)) + 1;
    const indptr = Uint32Array.from(data.indptr), indices = Uint32Array.from(data.indices);
    const E = indices.length;
    this.E = E;
    // the minSyn / sensory gate folded into the delivered counts (src/lifdiff.js:167-172)
    const sens = opts.sensoryMask || null;
    const counts = new Float32Array(E);
    for (let j = 0; j < E; j++) {
      const c = data.weights[j];
      counts[j] = (c >= p.minSyn && !(sens && sens[indices[j]])) ? c : 0;
    }
    const sign = opts.preSign ? Float32Array.from(opts.preSign)
      : Float32Array.from({ length: N }, (_, i) => EXC_SIGN[data.nt[i]]);
    this._h = native.diffCreate(N, indptr, indices, counts, sign, diffParamsVec(p));
    for (const k of ['sizeLog', 'logGain', 'outScale', 'thrOffset', 'inScale']) {
      if (opts[k]) native.diffSetInput(this._h, INPUT[k], Float32Array.from(opts[k]));
    }
    if (typeof p.surrogateBeta !== 'number') native.diffSetInput(this._h, INPUT.surrogateBeta, Float32Array.from(p.surrogateBeta));
    if (opts.edgeGain) native.diffSetEdgeGain(this._h, Float32Array.from(opts.edgeGain));
    native.diffSetMasks(this._h, opts.thrMask ? Uint8Array.from(opts.thrMask) : new Uint8Array(N),
      opts.biasMask ? Uint8Array.from(opts.biasMask) : new Uint8Array(N));
    this.drive = new Float32Array(N);
    this.drivenSet = new Set();
    this.driveIdx = opts.driveIdx ? Int32Array.from(opts.driveIdx) : null;
    this.driveEpoch = Math.max(1, opts.driveEpoch || 1);
    this._tape = null;
    this._lossPer = null;
  }

  setDrive(list, rate) { for (let k = 0; k < list.length; k++) this.setDriveOne(list[k], rate); }
  setDriveOne(i, rate) { this.drive[i] = rate; if (rate > 0) this.drivenSet.add(i); else this.drivenSet.delete(i); }
  reset() {}                       // every forward pass starts from reset (src/lifdiff.js:222)

  _array(which) { const out = new Float32Array(this.N); native.diffGetState(this._h, STATE[which], out); return out; }
  get v() { return this._array('v'); }
  get gE() { return this._array('gE'); }
  get gI() { return this._array('gI'); }
  get adapt() { return this._array('adapt'); }
  get res() { return this._array('res'); }
  get refr() { return this._array('refr'); }
  get spikeCount() { return this._array('spikeCount'); }
  get inScale() { return this._array('inScale'); }
  get rng() { return native.diffRng(this._h); }

  // the drive and driven set, and the drive schedule when driveIdx, onEpoch or
  // driveSoft asks for one
  _uploadDrive(steps, onEpoch) {
    native.diffSetInput(this._h, INPUT.drive, this.drive);
    native.diffSetDriven(this._h, Int32Array.from(this.drivenSet));
    const ep = this.driveEpoch;
    if (onEpoch || this.driveIdx) {
      if (!this.driveIdx) throw new Error('LIFDiffHip: onEpoch needs driveIdx (the neurons whose drive it sets)');
      const idx = this.driveIdx, nE = Math.max(1, Math.ceil(steps / ep));
      const sched = new Float32Array(nE * idx.length);
      for (let e = 0; e < nE; e++) {
        if (onEpoch) onEpoch(e, e * ep);
        for (let k = 0; k < idx.length; k++) {
          const r = this.drive[idx[k]];
          // a positive rate that binary32 holds as 0 still joins the set (-0)
          sched[e * idx.length + k] = (r === 0 && this.drivenSet.has(idx[k])) ? -0 : r;
        }
      }
      native.diffSetDriveSchedule(this._h, idx, ep, sched, nE, this.p.driveSoft ? 1 : 0);
    } else if (this.p.driveSoft) {
      const idx = Int32Array.from(this.drivenSet), sched = new Float32Array(idx.length);
      for (let k = 0; k < idx.length; k++) sched[k] = this.drive[idx[k]];
      native.diffSetDriveSchedule(this._h, idx, Math.max(1, steps), sched, 1, 1);
    } else {
      native.diffSetDriveSchedule(this._h, null, 1, null, 0, 0);
    }
  }

  /** forward(steps, {seed, record, onEpoch, spikeTrain}) -> tape (src/lifdiff.js:219) */
  forward(steps, o) {
    o = o || {};
    const seed = o.seed === undefined ? 1 : o.seed, record = o.record === undefined ? true : o.record;
    this._uploadDrive(steps, o.onEpoch || null);
    this._uploadTrain(o.spikeTrain || null);
    const loss = native.diffForward(this._h, steps, seed, record ? 1 : 0);
    const tape = { steps, seed, loss, net: this, recorded: !!record };
    if (record && o.spikeTrain) tape.spikeTrain = o.spikeTrain;   // src/lifdiff.js:243
    this._tape = tape;
    return tape;
  }

  // spikeTrain (src/lifdiff.js:298-307) as offsets and indices; null clears
  _uploadTrain(train) {
    if (!train) { native.diffSetSpikeTrain(this._h, null, null); return; }
    const ptr = new Float64Array(train.length + 1);
    for (let t = 0; t < train.length; t++) ptr[t + 1] = ptr[t] + (train[t] ? train[t].length : 0);
    const idx = new Int32Array(Math.max(1, ptr[train.length]));
    let k = 0;
    for (let t = 0; t < train.length; t++) if (train[t]) for (const i of train[t]) idx[k++] = i;
    native.diffSetSpikeTrain(this._h, ptr, idx);
  }

  /** the tape's spike lists: [{idx, amp, forced}] per step, lif.c's order */
  spikes(tape) {
    const t = tape || this._tape;
    const off = new Float64Array(t.steps + 1);
    const total = native.diffSpikes(this._h, off, null, null);
    const idx = new Int32Array(Math.max(1, total)), amp = new Float32Array(Math.max(1, total));
    native.diffSpikes(this._h, off, idx, amp);
    const out = [];
    for (let s = 0; s < t.steps; s++) {
      // lif.c's order: forced spikes as the tape holds them (driven-set
      // order), then threshold spikes by index (the device appends those in
      // no particular order)
      const forced = [], thr = [];
      for (let k = off[s]; k < off[s + 1]; k++) {
        const e = { i: idx[k] & 0x7fffffff, a: amp[k] };
        if ((idx[k] & FORCED) !== 0) forced.push(e); else thr.push(e);
      }
      thr.sort((x, y) => x.i - y.i);
      const all = forced.concat(thr);
      out.push({ idx: Int32Array.from(all, (e) => e.i), amp: Float32Array.from(all, (e) => e.a),
        forced: all.map((e, k) => k < forced.length) });
    }
    return out;
  }

  /**
   * backward(tape, dLdSpike, {truncate, lossFrom, dLdSpikePerEpoch, edges})
   * -> {params, logGain, drive, truncate, clamped, edge?, inScale}
   * (src/lifdiff.js:494-686; edge = dL/d edgeGain when `edges`)
   */
  backward(tape, dLdSpike, o) {
    o = o || {};
    if (!tape || tape !== this._tape || !tape.recorded) throw new Error('LIFDiffHip: backward needs the tape of this network\'s last recorded forward pass');
    const truncate = o.truncate || 0;
    native.diffSetLoss(this._h, dLdSpike ? Float32Array.from(dLdSpike) : null, o.lossFrom || 0);
    if (o.dLdSpikePerEpoch) {
      const nE = Math.floor(o.dLdSpikePerEpoch.length / this.N);
      native.diffSetLossPerEpoch(this._h, Float32Array.from(o.dLdSpikePerEpoch), nE, this.driveEpoch);
      this._lossPer = true;
    } else if (this._lossPer) {
      native.diffSetLossPerEpoch(this._h, null, 1, this.driveEpoch);
      this._lossPer = false;
    }
    const pp = new Float64Array(9), lg = new Float32Array(this.N), gi = new Float32Array(this.N);
    const ge = o.edges ? new Float32Array(this.E) : null, cl = new Int32Array(3);
    native.diffBackward(this._h, truncate, pp, lg, ge, gi, cl);
    return this._result(pp, lg, ge, gi, cl, truncate, tape.steps);
  }

  _result(pp, lg, ge, gi, cl, truncate, steps) {
    const params = {};
    DIFF_PARAMS.forEach((k, i) => { params[k] = pp[i]; });
    let drive = null;
    if (this.driveIdx) {
      const nE = Math.max(1, Math.ceil(steps / this.driveEpoch));
      drive = new Float32Array(nE * this.driveIdx.length);
      native.diffDriveGrad(this._h, drive, drive.length);
    }
    const out = { params, logGain: lg, drive, truncate, inScale: gi,
      clamped: { params: cl[0], logGain: cl[1], edge: cl[2] } };
    if (ge) out.edge = ge;
    return out;
  }

  /** dL/d gain_T for outScale = gain of the neuron's type (src/brainmodel.js:63-100) */
  typeGrads(typeOf, nTypes) {
    const out = new Float64Array(nTypes);
    native.diffTypeGrads(this._h, Int32Array.from(typeOf), nTypes, out);
    return out;
  }

  stat(name) { return native.diffStat(this._h, name); }
  destroy() { if (this._h) { native.diffDestroy(this._h); this._h = null; } }
}

/**
 * LIFDiffBatchHip: B LIFDiffHip networks over one graph, each with its own
 * options and drive, stepped in lockstep (one launch per step kernel for all
 * of them), as the many-instance forward + adjoint of a fit.  No counterpart
 * in the reference's API (it runs one network per worker).
 */
class LIFDiffBatchHip {
  constructor(members) {
    this.members = members;
    this._h = native.diffBatchCreate(members.map((m) => m._h));
  }

  /** forward(steps[B], seeds[B]) -> tapes */
  forward(steps, seeds) {
    const B = this.members.length;
    this.members.forEach((m, b) => m._uploadDrive(steps[b], null));
    const L = new Float64Array(B);
    native.diffBatchForward(this._h, Int32Array.from(steps), Float64Array.from(seeds), 1, L);
    return this.members.map((m, b) => {
      const tape = { steps: steps[b], seed: seeds[b], loss: L[b], net: m, recorded: true };
      m._tape = tape;
      return tape;
    });
  }

  /** backward(dLdSpike[B] (each Float32Array or null), {truncate, edges}) -> results */
  backward(dLdSpike, o) {
    o = o || {};
    this.members.forEach((m, b) => native.diffSetLoss(m._h, dLdSpike && dLdSpike[b] ? Float32Array.from(dLdSpike[b]) : null, 0));
    native.diffBatchBackward(this._h, o.truncate || 0);
    return this.members.map((m) => {
      const pp = new Float64Array(9), lg = new Float32Array(m.N), gi = new Float32Array(m.N);
      const ge = o.edges ? new Float32Array(m.E) : null, cl = new Int32Array(3);
      native.diffBackward(m._h, -1, pp, lg, ge, gi, cl);
      return m._result(pp, lg, ge, gi, cl, o.truncate || 0, m._tape.steps);
    });
  }

  /** the members' type-gain gradients summed (each with its own gains) */
  typeGrads(typeOf, nTypes) {
    const out = new Float64Array(nTypes);
    for (const m of this.members) { const g = m.typeGrads(typeOf, nTypes); for (let t = 0; t < nTypes; t++) out[t] += g[t]; }
    return out;
  }

  destroy() {
    if (this._h) { native.diffBatchDestroy(this._h); this._h = null; }
    for (const m of this.members) m.destroy();
  }
}

module.exports = { LIFDiffHip, LIFDiffBatchHip, DIFF_DEFAULTS, DIFF_PARAMS };
