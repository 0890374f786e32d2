'use strict';
// Reads that resolve later, from JavaScript: node async_read.js <outdir>
// 1. P.read snapshots in call order: a read issued between two steps holds the state after the
//    first, whatever is enqueued before it is awaited (with sortEvery changed to 2 before stepping).
// 2. P2G into a second grid array with the global path equals the tiled P2G into the simulation's
//    grid (physics.test.js:33-42 through the JS host).
// 3. A pending P.read does not block the JS thread: with about a thousand substeps queued ahead of
//    it, a setImmediate callback queued after the read runs before the read resolves.
const fs = require('fs');
const path = require('path');
const SI = require('../../shaderintel_amd/js/index.js');

async function main() {
  const outdir = process.argv[2];
  const gpu = await SI.init();
  const summary = {};

  const sim = gpu.mpm({ grid: 32, maxParticles: 40000, substeps: 8 });
  sim.add({ box: [[0.2, 0.15, 0.2], [0.5, 0.55, 0.45]] }, { material: 'water' });
  sim.add({ sphere: [[0.7, 0.4, 0.5], 0.12] }, { material: 'jelly', velocity: [-0.3, 0, 0] });
  const n = sim.count;
  const before = await sim.P.read(n * 8);
  sim.sortEvery = 2;
  sim.step();
  const pending = sim.P.read(n * 8);
  sim.step();
  const after = await sim.P.read(n * 8);
  const mid = await pending;
  for (const [name, a] of [['before', before], ['mid', mid], ['after', after]])
    fs.writeFileSync(path.join(outdir, name + '.f32'), Buffer.from(a.buffer));
  fs.writeFileSync(path.join(outdir, 'materials.f32'), Buffer.from(sim.materialTable().buffer));
  Object.assign(summary, { count: n, dims: sim.dims, gravity: Array.from(sim.latticeGravity()), sortEvery: sim.sortEvery });

  sim._sort();
  const other = gpu.array('vec4i', sim.numNodes);
  sim.p2g(sim.P, true, other);
  sim.p2g(sim.P, false, sim.grid);
  const a = await other.read(), b = await sim.grid.read();
  let differ = 0, nonzero = 0;
  for (let i = 0; i < a.length; i++) { if (a[i] !== b[i]) differ++; if (a[i] !== 0) nonzero++; }
  Object.assign(summary, { p2gDiffer: differ, p2gNonzero: nonzero, p2gWords: a.length });
  let bad = null;
  try { sim.p2g(sim.P, true, gpu.array('vec4i', 5)); } catch (e) { bad = e.code; }
  summary.badGrid = bad;
  sim.destroy();

  const big = gpu.mpm({ grid: 64, maxParticles: 800000, substeps: 1024 });
  big.add({ box: [[0.05, 0.05, 0.05], [0.95, 0.5, 0.95]] }, { material: 'water' });
  big.step();
  const order = [];
  const t0 = process.hrtime.bigint();
  const read = big.P.read(big.count * 8).then((r) => { order.push('read'); return r; });
  const tCall = Number(process.hrtime.bigint() - t0) / 1e6;
  setImmediate(() => order.push('timer'));
  const r = await read;
  const tResolve = Number(process.hrtime.bigint() - t0) / 1e6;
  let finite = 0;
  for (let k = 0; k < big.count; k++) if (Number.isFinite(r[32 * k]) && Number.isFinite(r[32 * k + 1])) finite++;
  Object.assign(summary, { order, tCallMs: tCall, tResolveMs: tResolve, bigCount: big.count, bigFinite: finite });
  big.destroy();
  gpu.destroy();
  fs.writeFileSync(path.join(outdir, 'summary.json'), JSON.stringify(summary));
  console.log('async read done', JSON.stringify(summary));
}
main().catch((e) => { console.error('FAILED', e && e.stack || e); process.exit(1); });
