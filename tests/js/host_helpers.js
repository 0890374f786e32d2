'use strict';
// Writes the JS host's material table, shape samples and packed records to a JSON file so the
// Python test can compare them with the oracle's (no device needed).
const fs = require('fs');
const SI = require('../../shaderintel_amd/js/index.js');
const out = process.argv[2];
const N = SI.native();
const res = { materials: {}, samples: {}, packed: null, sdf: null };
for (const name of Object.keys(SI.MATERIALS)) {
  const m = SI.MATERIALS[name];
  const row = N.materialRow(m.kind, m.density === undefined ? 1 : m.density, m.stiffness === undefined ? 1 : m.stiffness,
    m.poisson === undefined ? 0.3 : m.poisson, m.viscosity === undefined ? 0 : m.viscosity, m.hardening === undefined ? 10 : m.hardening,
    m.compression === undefined ? 0.025 : m.compression, m.stretch === undefined ? 0.0075 : m.stretch, m.friction === undefined ? 33 : m.friction, 2);
  res.materials[name] = Array.from(new Uint32Array(row.buffer));
}
const box = SI.sampleShape({ box: [[0.1, 0.2, 0.3], [0.4, 0.35, 0.5]] }, 1 / 128, 0.6, 7);
const sph = SI.sampleShape({ sphere: [[0.5, 0.5, 0.5], 0.15] }, 1 / 96, 1.0, 3);
const cyl = SI.sampleShape({ cylinder: [[0.5, 0.4, 0.5], 0.1, 0.2] }, 1 / 96, 0.6, 11);
const sdf = SI.sampleShape({ sdf: (p) => Math.hypot(p[0] - 0.5, p[1] - 0.5, p[2] - 0.5) - 0.15, bounds: [[0.35, 0.35, 0.35], [0.65, 0.65, 0.65]] }, 1 / 96, 1.0, 3);
res.samples = { box: Array.from(new Uint32Array(box.buffer)), sphere: Array.from(new Uint32Array(sph.buffer)),
  cylinder: Array.from(new Uint32Array(cyl.buffer)), sdf: Array.from(new Uint32Array(sdf.buffer)) };
const rec = N.packParticles(box, Float64Array.from([0, 0, 0]), 1 / 64, Int32Array.from([64, 64, 64]), 3, 2,
  Float64Array.from([0.1, -0.2, 0.05]), (1 / 480) / (1 / 64), 0.01, 12345);
res.packed = Array.from(new Uint32Array(rec.buffer));
let code = null;
if (process.argv[3] === 'nodevice') {
  try { new SI.Context(0); } catch (e) { code = e.code + '|' + (e instanceof SI.SIError); }
}
res.nodevice = code;
fs.writeFileSync(out, JSON.stringify(res));
