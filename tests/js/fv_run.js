'use strict';
// Runs two eyes of the optic lobe through the addon's FlyVisDiffHip for the
// parity test: reads a JSON case (model, luminance frames per eye, dLdV per
// eye) and writes the final activity and the shared gradients as JSON.
const fs = require('fs');
const path = require('path');
const { FlyVisDiffHip } = require(path.join(__dirname, '..', '..', 'fly-brain_amd', 'js', 'flyviship.js'));

const spec = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
const eyes = [new FlyVisDiffHip(spec.model)];
eyes.push(new FlyVisDiffHip(null, eyes[0].sharedParts));
const T = spec.lums[0].length;
for (let e = 0; e < 2; e++) {
  eyes[e].startTape(T);
  for (let t = 0; t < T; t++) { eyes[e].setInput(Float32Array.from(spec.lums[e][t])); eyes[e].stepTaped(); }
}
let g = null;
const lum = [];
for (let e = 0; e < 2; e++) {
  const dl = new Float32Array(T * eyes[e].nCol);
  g = eyes[e].backward(Float32Array.from(spec.dLdV[e]), g, dl);
  lum.push(Array.from(dl));
}
const out = { v: eyes.map(e => Array.from(e.v)), weight: Array.from(g.weight), bias: Array.from(g.bias),
  kdt: Array.from(g.kdt), lum, tape: eyes[0].tapeLength };
eyes[1].destroy(); eyes[0].destroy();
process.stdout.write(JSON.stringify(out));
