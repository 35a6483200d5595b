'use strict';
// Runs the differentiable model or an ensemble through the addon's
// LIFDiffHip / LIFDiffBatchHip / LIFBatchHip classes for the parity tests:
// reads a JSON case and writes the results as JSON.
//   mode 'diff':      {data, opts, drive: [[i, hz]...], steps, seed, dL, edges, spikeTrain}
//                     -> {loss, params, logGain, clamped, spikes, v, rng, trainKept}
//   mode 'diffbatch': {data, members: [{opts, drive, dL}], steps[], seeds[]}
//                     -> per member {loss, params, logGain} batched and alone
//   mode 'batch':     {graph, N, params, seeds, typeOf, gains, drive: [ix, hz],
//                     silence: [[b, ix]...], steps[], run}
//                     -> per member {spikeCount, v, rng}
const fs = require('fs');
const path = require('path');
const js = path.join(__dirname, '..', '..', 'fly-brain_amd', 'js');
const { LIFDiffHip, LIFDiffBatchHip } = require(path.join(js, 'lifdiffhip.js'));
const { LIFBatchHip } = require(path.join(js, 'lifbatchhip.js'));

const spec = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));

function build(data, opts, drive) {
  const o = Object.assign({}, opts);
  for (const k of ['sizeLog', 'logGain', 'outScale', 'thrOffset', 'inScale', 'edgeGain']) if (o[k]) o[k] = Float32Array.from(o[k]);
  for (const k of ['thrMask', 'biasMask']) if (o[k]) o[k] = Uint8Array.from(o[k]);
  if (Array.isArray(o.surrogateBeta)) o.surrogateBeta = Float32Array.from(o.surrogateBeta);
  const net = new LIFDiffHip(data, o);
  for (const [i, hz] of drive || []) net.setDriveOne(i, hz);
  return net;
}

// the state the forward pass left (read before the backward pass, whose
// segment replays move the live state)
function afterForward(net, tape) {
  const sp = net.spikes(tape);
  const idx = [], off = [0];
  for (const s of sp) { for (let k = 0; k < s.idx.length; k++) idx.push(s.idx[k]); off.push(idx.length); }
  return { spikeIdx: idx, spikeOff: off, v: Array.from(net.v), rng: net.rng, spikeCount: Array.from(net.spikeCount) };
}

function result(fwd, tape, g) {
  return Object.assign({ loss: tape.loss, params: g.params, logGain: Array.from(g.logGain), clamped: g.clamped,
    edge: g.edge ? Array.from(g.edge) : null }, fwd);
}

let out;
if (spec.mode === 'diff') {
  const net = build(spec.data, spec.opts, spec.drive);
  const train = spec.spikeTrain || null;
  const tape = net.forward(spec.steps, { seed: spec.seed, spikeTrain: train });
  const fwd = afterForward(net, tape);
  const g = net.backward(tape, Float32Array.from(spec.dL), { edges: !!spec.edges });
  out = result(fwd, tape, g);
  if (train) out.trainKept = tape.spikeTrain === train;
  net.destroy();
} else if (spec.mode === 'diffbatch') {
  const ms = spec.members.map((m) => build(spec.data, m.opts, m.drive));
  const bt = new LIFDiffBatchHip(ms);
  const tapes = bt.forward(spec.steps, spec.seeds);
  const fwds = ms.map((m, b) => afterForward(m, tapes[b]));
  const gs = bt.backward(spec.members.map((m) => Float32Array.from(m.dL)), { edges: true });
  const batched = ms.map((m, b) => result(fwds[b], tapes[b], gs[b]));
  bt.destroy();
  const alone = spec.members.map((m, b) => {
    const net = build(spec.data, m.opts, m.drive);
    const tape = net.forward(spec.steps[b], { seed: spec.seeds[b] });
    const fwd = afterForward(net, tape);
    const r = result(fwd, tape, net.backward(tape, Float32Array.from(m.dL), { edges: true }));
    net.destroy();
    return r;
  });
  out = { batched, alone };
} else if (spec.mode === 'batch') {
  const bt = new LIFBatchHip({ graph: spec.graph, N: spec.N, params: spec.params, seeds: spec.seeds,
    typeOf: spec.typeOf, gains: spec.gains });
  bt.setDrive(spec.drive[0], spec.drive[1]);
  for (const [b, ix] of spec.silence) bt.silence(b, ix);
  bt.setSteps(spec.steps);
  bt.run(spec.run);
  out = { members: [] };
  const rng = bt.rng;
  for (let b = 0; b < bt.B; b++) out.members.push({ spikeCount: Array.from(bt.spikeCount(b)), v: Array.from(bt.array('v', b)), rng: rng[b] });
  out.counters = bt.counters();
  bt.destroy();
} else {
  throw new Error('unknown mode ' + spec.mode);
}
process.stdout.write(JSON.stringify(out));
