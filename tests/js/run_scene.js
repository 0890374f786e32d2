'use strict';
// Runs a scene through the JS host on device 0 and writes particles (float32 records) and the grid
// words to binary files, plus a JSON summary: node run_scene.js <outdir> <frames>
const fs = require('fs');
const path = require('path');
const SI = require('../../shaderintel_amd/js/index.js');

async function main() {
  const outdir = process.argv[2];
  const frames = +(process.argv[3] || 1);
  const gpu = await SI.init();
  const sim = gpu.mpm({ grid: 32, maxParticles: 40000, substeps: 8 });
  sim.add({ box: [[0.2, 0.15, 0.2], [0.5, 0.55, 0.45]] }, { material: 'water', jitter: 1.0 });
  sim.add({ sphere: [[0.7, 0.4, 0.5], 0.12] }, { material: 'jelly', velocity: [-0.3, 0, 0] });
  sim.add({ box: [[0.55, 0.6, 0.55], [0.75, 0.7, 0.75]] }, { material: 'snow' });
  sim.add({ cylinder: [[0.3, 0.75, 0.7], 0.08, 0.06] }, { material: 'sand' });
  const before = await sim.P.read(sim.count * 8);
  fs.writeFileSync(path.join(outdir, 'before.f32'), Buffer.from(before.buffer));
  fs.writeFileSync(path.join(outdir, 'materials.f32'), Buffer.from(sim.materialTable().buffer));
  for (let f = 0; f < frames; f++) sim.step();
  const after = await sim.P.read(sim.count * 8);
  fs.writeFileSync(path.join(outdir, 'after.f32'), Buffer.from(after.buffer));
  const grid = await sim.grid.read();
  fs.writeFileSync(path.join(outdir, 'grid.i32'), Buffer.from(grid.buffer));
  const pos = await sim.readPositions();
  fs.writeFileSync(path.join(outdir, 'positions.f32'), Buffer.from(pos.buffer));
  // one more substep split at the coupling call site, with the grid read back and written unchanged
  sim.substepBegin();
  const mid = await sim.grid.read();
  sim.grid.write(mid);
  sim.substepEnd();
  const after2 = await sim.P.read(sim.count * 8);
  fs.writeFileSync(path.join(outdir, 'after2.f32'), Buffer.from(after2.buffer));
  const law = await SI.checkMPM(sim);
  const summary = { count: sim.count, dims: sim.dims, stepCount: sim.stepCount, gravity: Array.from(sim.latticeGravity()), law,
    dx: sim.dx };
  sim.destroy(); sim.destroy();
  let threw = null;
  try { sim.step(); } catch (e) { threw = e.code; }
  summary.afterDestroy = threw;
  gpu.destroy();
  fs.writeFileSync(path.join(outdir, 'summary.json'), JSON.stringify(summary));
  console.log('js scene done', JSON.stringify(summary));
}
main().catch((e) => { console.error('FAILED', e && e.stack || e); process.exit(1); });
