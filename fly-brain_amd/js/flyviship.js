'use strict';
// The optic lobe on the GPU for Node: FlyVis (src/flyvis.js:3-25) and
// FlyVisDiff (src/flyvisdiff.js:36-185) over the addon's fv* functions
// (fb_fv_* of include/flybrain.h).  A model is parseFlyVis's object
// ({N, E, bias, tau, indptr, target, weight, dt, inputIdx, kdt?}); the
// second eye is built on the first eye's `sharedParts`.  Activity and
// gradients are bit-identical to the reference's arithmetic.
const { native } = require('./lifhip.js');

class FlyVisModelHip {
  constructor(model) {
    this.N = model.N; this.E = model.E; this.dt = model.dt || 0.02;
    this.h = native.fvModelCreate(model.N, Int32Array.from(model.indptr), Int32Array.from(model.target),
      Float32Array.from(model.weight), Float32Array.from(model.bias), Float32Array.from(model.tau), this.dt);
    if (model.kdt) native.fvModelSetKdt(this.h, Float32Array.from(model.kdt));
    // the photoreceptor map, flattened type by type (src/flyvisdiff.js:60-64)
    const ix = [], col = [];
    for (const t in (model.inputIdx || {})) { const a = model.inputIdx[t];
      for (let k = 0; k < a.length; k++) { ix.push(a[k]); col.push(k); } }
    this.nCol = col.length ? Math.max(...col) + 1 : 0;
    if (ix.length) native.fvModelSetInputMap(this.h, Int32Array.from(ix), Int32Array.from(col));
    this.refs = 1;
  }
  retain() { this.refs++; return this; }
  release() { if (--this.refs === 0 && this.h) { native.fvModelRelease(this.h); this.h = null; } }
}

class FlyVisDiffHip {
  /** new FlyVisDiff(model, shared): `shared` is another instance's sharedParts */
  constructor(model, shared = null) {
    this.model = shared ? shared.retain() : new FlyVisModelHip(model);
    this.N = this.model.N; this.E = this.model.E; this.nCol = this.model.nCol; this.dt = this.model.dt;
    this.h = native.fvCreate(this.model.h);
    this.tape = null;
  }
  get sharedParts() { return this.model; }
  reset() { native.fvReset(this.h); }
  /** lum: Float32Array(nCol), luminance per hex column */
  setInput(lum) { native.fvSetLum(this.h, lum instanceof Float32Array ? lum : Float32Array.from(lum)); }
  step() { native.fvStep(this.h, 1); }
  settle(lum, steps = 150) { this.reset(); this.setInput(lum); native.fvStep(this.h, steps); }
  startTape(steps) { native.fvTapeStart(this.h, steps); this.tape = { steps }; return this.tape; }
  stepTaped() { native.fvStepTaped(this.h, 1); }
  get tapeLength() { return native.fvTapeLen(this.h); }
  /** backward(dLdV, g, dLdLum) (src/flyvisdiff.js:128-185): g = {weight, bias, kdt}
   *  Float32Array accumulators, added to (created when absent); dLdLum optional */
  backward(dLdV, g = null, dLdLum = null) {
    g = g || { weight: new Float32Array(this.E), bias: new Float32Array(this.N), kdt: new Float32Array(this.N) };
    native.fvBackward(this.h, dLdV, g.weight, g.bias, g.kdt, dLdLum);
    return g;
  }
  get v() { const out = new Float32Array(this.N); native.fvGet(this.h, 0, out); return out; }
  set v(x) { native.fvSetV(this.h, Float32Array.from(x)); }
  get acc() { const out = new Float32Array(this.N); native.fvGet(this.h, 1, out); return out; }
  destroy() { if (this.h) { native.fvDestroy(this.h); this.h = null; this.model.release(); } }
}

/** FlyVis (src/flyvis.js): the forward half */
class FlyVisHip extends FlyVisDiffHip {}

module.exports = { FlyVisHip, FlyVisDiffHip, FlyVisModelHip };
