'use strict';
// Short TypedArrays handed to the addon must raise a RangeError before any
// copy: a brain (N = 64) and an optic-lobe instance (N = 8, 10 edges, 3
// luminance columns) are created, and every array-taking call is tried with
// an array one element short.  Writes {call: "RangeError" | other} as JSON.
const path = require('path');
const { native } = require(path.join(__dirname, '..', '..', 'fly-brain_amd', 'js', 'lifhip.js'));

const out = {};
function probe(name, f) {
  try { f(); out[name] = 'no error'; } catch (e) { out[name] = e instanceof RangeError ? 'RangeError' : String(e); }
}
const N = 64;
const indptr = new Uint32Array(N + 1);
for (let i = 0; i <= N; i++) indptr[i] = i;
const indices = new Uint32Array(N);
for (let i = 0; i < N; i++) indices[i] = (i + 1) % N;
const { LIFHip } = require(path.join(__dirname, '..', '..', 'fly-brain_amd', 'js', 'lifhip.js'));
const brain = new LIFHip({ graph: { indptr: Array.from(indptr), indices: Array.from(indices),
  weights: new Array(N).fill(1), sign: new Array(N).fill(0.5) }, N, seed: 1 });
probe('getArray', () => native.getArray(brain._h, 0, new Float32Array(N - 1)));
probe('setArray', () => native.setArray(brain._h, 0, new Float32Array(N - 1)));
probe('getArrayOk', () => native.getArray(brain._h, 0, new Float32Array(N)));
brain.destroy();

const fN = 8, fE = 10;
const fip = Int32Array.from([0, 2, 3, 5, 6, 7, 8, 9, 10]);
const ftg = Int32Array.from([1, 2, 3, 4, 5, 6, 7, 0, 1, 2]);
const m = native.fvModelCreate(fN, fip, ftg, new Float32Array(fE).fill(0.1), new Float32Array(fN).fill(0.2),
  new Float32Array(fN).fill(5), 1.0);
native.fvModelSetInputMap(m, Int32Array.from([0, 1, 2]), Int32Array.from([0, 1, 2]));
probe('fvModelSetKdt', () => native.fvModelSetKdt(m, new Float32Array(fN - 1)));
const e = native.fvCreate(m);
probe('fvSetInput', () => native.fvSetInput(e, new Float32Array(fN - 1)));
probe('fvSetV', () => native.fvSetV(e, new Float32Array(fN - 1)));
probe('fvSetLum', () => native.fvSetLum(e, new Float32Array(2)));
probe('fvGet', () => native.fvGet(e, 0, new Float32Array(fN - 1)));
native.fvTapeStart(e, 2);
native.fvSetLum(e, new Float32Array(3));
native.fvStepTaped(e, 2);
const full = () => [new Float32Array(2 * fN), new Float32Array(fE), new Float32Array(fN), new Float32Array(fN),
  new Float32Array(2 * 3)];
const short = (k) => { const a = full(); a[k] = new Float32Array(a[k].length - 1); return a; };
['dLdV', 'gWeight', 'gBias', 'gKdt', 'dLdLum'].forEach((nm, k) => {
  probe('fvBackward_' + nm, () => { const a = short(k); native.fvBackward(e, a[0], a[1], a[2], a[3], a[4]); });
});
probe('fvBackwardOk', () => { const a = full(); native.fvBackward(e, a[0], a[1], a[2], a[3], a[4]); });
native.fvDestroy(e);
native.fvModelRelease(m);
process.stdout.write(JSON.stringify(out));
