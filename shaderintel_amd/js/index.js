'use strict';
// JavaScript host of the MI355X MLS-MPM library: ShaderIntel's MPM world API (src/physics/mpm.js:82-680,
// types/shaderintel.d.ts:813-900) over the C ABI of libmpm_hip.so, reached through the N-API addon.
//
// Host-side math (material table, gravity, unit conversions, shape sampling) is done in float64 like
// the reference, so the values handed to the device are the same float32 bits the reference would
// upload. Reads return Promises like the reference's GArray.read (they resolve once the device copy
// is done). Written for Node >= 12 (no ?? / ?. / ||=).
const path = require('path');

let addon = null;
function native() {
  if (!addon) addon = require(path.join(__dirname, 'mpm_napi.node'));
  return addon;
}

class SIError extends Error {   // src/core/errors.js:5-11
  constructor(msg, code) {
    super(String(msg).startsWith('ShaderIntel') ? msg : 'ShaderIntel: ' + msg);
    this.name = 'ShaderIntelError';
    this.code = code || 'error';
  }
}

function call(fn) {             // rethrow addon errors as SIError with the same code
  try { return fn(); } catch (e) {
    if (e instanceof SIError) throw e;
    throw new SIError(e.message, e.code || 'error');
  }
}

function callAsync(fn) {        // a native read that resolves later: both kinds of failure as SIError
  return call(fn).catch((e) => { throw e instanceof SIError ? e : new SIError(e.message, e.code || 'error'); });
}

const FLUID = 0, JELLY = 1, SNOW = 2, SAND = 3;
const MPM_KINDS = { fluid: FLUID, jelly: JELLY, elastic: JELLY, snow: SNOW, sand: SAND };
const MAX_DENSITY = 50;
const MATERIALS = {             // mpm.js:38-46
  water: { kind: FLUID, density: 1, stiffness: 1.0, viscosity: 0.0, color: [0.12, 0.42, 0.95] },
  oil: { kind: FLUID, density: 0.9, stiffness: 0.9, viscosity: 0.3, color: [0.9, 0.62, 0.12] },
  honey: { kind: FLUID, density: 1.4, stiffness: 0.9, viscosity: 1.0, color: [0.95, 0.6, 0.05] },
  jelly: { kind: JELLY, density: 1, stiffness: 0.35, poisson: 0.3, color: [0.95, 0.3, 0.35] },
  rubber: { kind: JELLY, density: 1.2, stiffness: 0.7, poisson: 0.45, color: [0.25, 0.25, 0.28] },
  snow: { kind: SNOW, density: 0.4, stiffness: 0.6, poisson: 0.2, hardening: 10, compression: 0.025, stretch: 0.0075, color: [0.95, 0.96, 1.0] },
  sand: { kind: SAND, density: 1.6, stiffness: 0.6, poisson: 0.3, friction: 33, color: [0.86, 0.68, 0.42] },
};
const PARTICLE_FLOATS = 32;

function dflt(v, d) { return v === undefined || v === null ? d : v; }

/** A device (stands in for SI.init's Context, src/core/context.js:32-100). */
class Context {
  constructor(device) {
    this._h = call(() => native().ctxCreate(device || 0));
    const info = native().ctxInfo(this._h);
    this.name = info.name;
    this.cus = info.cus;
    this.destroyed = false;
  }
  mpm(o) { return new MPM(this, o); }                         // gpu.mpm(o), src/index.js:87
  /** gpu.array(type, n) for the grids MPM.p2g can target ('vec4i', n nodes); read() resolves later. */
  array(type, n) {
    if (type !== 'vec4i') throw new SIError(`gpu.array: only 'vec4i' grids are supported here (got ${type})`, 'unsupported');
    const data = new Int32Array(n * 4);
    return { type, length: n, data, read() { return Promise.resolve(data.slice()); } };
  }
  /** Scan.run (src/prims/scan.js:292-306): exclusive prefix sum of a Uint32Array. */
  scan(a) { return call(() => native().scanU32(this._h, a)); }
  /** SpatialGrid.build (src/prims/grid.js:70-79): { order, start } of a counting sort by key. */
  countingSort(keys, numBuckets) { return call(() => native().countingSort(this._h, keys, numBuckets)); }
  destroy() {
    if (this.destroyed) return;
    this.destroyed = true;
    native().ctxDestroy(this._h);
  }
}

async function init(o) { return new Context(o && o.device); }

class MPM {
  constructor(ctx, o) {
    o = o || {};
    this.ctx = ctx;
    this.maxParticles = o.maxParticles || 262144;
    const size = Array.isArray(o.size) ? o.size : [o.size || 1, o.size || 1, o.size || 1];
    const res = o.grid || 64;
    this.dx = Math.max.apply(null, size) / res;
    this.dims = size.map((s) => Math.max(8, Math.round(s / this.dx)));
    this.size = this.dims.map((d) => d * this.dx);
    this.origin = o.origin || [0, 0, 0];
    this.gravity = o.gravity || [0, -9.81, 0];
    this._substeps = Math.max(1, Math.round(o.substeps || 8));
    this._timeScale = dflt(o.timeScale, 1);
    this._sortEvery = o.sortEvery || 4;
    this.ppcAxis = o.ppcAxis || 2;
    this.bound = 3;
    this.count = 0;
    this.materials = [];
    this.materialIndex = new Map();
    this.colliders = [];
    this.numNodes = this.dims[0] * this.dims[1] * this.dims[2];
    this.destroyed = false;
    this._h = call(() => native().simCreate(ctx._h, { dims: this.dims, maxParticles: this.maxParticles,
      ppcAxis: this.ppcAxis, bound: this.bound, sortEvery: this._sortEvery, flags: 0 }));
    const sim = this;
    // GArray-like views of the particle buffer and the grid (offsets and counts in vec4 units).
    this.P = {
      read(n4) { return callAsync(() => native().readParticles(sim._alive('MPM.P.read'), 0, Math.floor((n4 === undefined ? sim.count * 8 : n4) / 8))); },
      write(data, offset4) { call(() => native().writeParticles(sim._alive('MPM.P.write'), data, Math.floor((offset4 || 0) / 8))); },
      get destroyed() { return sim.destroyed; },
    };
    this.grid = {
      read() { return callAsync(() => native().readGrid(sim._alive('MPM.grid.read'), sim.numNodes)); },
      /** Between substepBegin and substepEnd: every node's f32 (v, m) bits, reference layout. */
      write(words) { call(() => native().writeGrid(sim._alive('MPM.grid.write'), words)); },
      get destroyed() { return sim.destroyed; },
    };
  }

  get stepCount() { return this.destroyed ? this._lastStepCount : native().getStepCount(this._h); }
  set stepCount(n) { native().setStepCount(this._alive('MPM.stepCount'), n); }
  get substepDt() { return (1 / 60) * this.timeScale / this.substeps; }      // mpm.js:130
  get substeps() { return this._substeps; }
  set substeps(n) { this._retime(() => { this._substeps = Math.max(1, Math.round(n)); }); }
  get timeScale() { return this._timeScale; }
  set timeScale(x) { this._retime(() => { this._timeScale = Math.max(1e-6, +x); }); }
  _retime(change) {                                                              // mpm.js:140-145
    const before = this.substepDt;
    change();
    const k = this.substepDt / before;
    if (this.count && !this.destroyed && Math.abs(k - 1) > 1e-12) call(() => native().rescale(this._h, k));
  }

  _alive(what) {
    if (this.destroyed) throw new SIError(what + ': it was destroyed', 'destroyed');
    return this._h;
  }

  // ---- materials (mpm.js:148-199) ----
  _material(nameOrDef) {
    const def = typeof nameOrDef === 'string' ? Object.assign({}, MATERIALS[nameOrDef], { name: nameOrDef }) : Object.assign({}, nameOrDef);
    if (!def || def.kind === undefined)
      throw new SIError('MPM: unknown material ' + JSON.stringify(nameOrDef) + ' (use ' + Object.keys(MATERIALS).join(', ') + ' or a definition with a kind)', 'invalid-argument');
    if (typeof def.kind === 'string') {
      if (!(def.kind in MPM_KINDS)) throw new SIError('MPM: unknown material kind "' + def.kind + '" (use ' + Object.keys(MPM_KINDS).join(', ') + ')', 'invalid-argument');
      def.kind = MPM_KINDS[def.kind];
    }
    if ([FLUID, JELLY, SNOW, SAND].indexOf(def.kind) < 0) throw new SIError('MPM: material kind must be 0–3 (got ' + def.kind + ')', 'invalid-argument');
    const density = dflt(def.density, 1);
    if (!(density > 0 && density <= MAX_DENSITY))
      throw new SIError('MPM: density ' + density + ' is outside (0, ' + MAX_DENSITY + '] — densities are relative to water (1), not kg/m³', 'invalid-argument');
    const key = JSON.stringify(def);
    if (this.materialIndex.has(key)) return this.materialIndex.get(key);
    if (this.materials.length >= 8) throw new SIError('MPM: at most 8 distinct materials at once (clear() frees them)', 'invalid-argument');
    const idx = this.materials.length;
    this.materials.push(def);
    this.materialIndex.set(key, idx);
    this._uploadMaterials();
    return idx;
  }

  materialTable() {
    const data = new Float32Array(8 * 12);
    this.materials.forEach((m, k) => {
      data.set(native().materialRow(m.kind, dflt(m.density, 1), dflt(m.stiffness, 1), dflt(m.poisson, 0.3), dflt(m.viscosity, 0),
        dflt(m.hardening, 10), dflt(m.compression, 0.025), dflt(m.stretch, 0.0075), dflt(m.friction, 33), this.ppcAxis), k * 12);
    });
    return data;
  }

  _uploadMaterials() { call(() => native().setMaterials(this._h, this.materialTable(), 8)); }

  // ---- emission (mpm.js:207-232) ----
  add(shape, opts) {
    opts = opts || {};
    this._alive('MPM.add');
    const mat = this._material(opts.material || shape.material || 'water');
    const vel = opts.velocity || shape.velocity || [0, 0, 0];
    const pts = shape.points ? Float32Array.from(shape.points) : sampleShape(shape, this.dx / this.ppcAxis, dflt(opts.jitter, 0.6), dflt(opts.seed, this.count + 1));
    const n = Math.min(pts.length / 3, this.maxParticles - this.count);
    if (n <= 0) return 0;
    const s = this.substepDt / this.dx;
    const noise = opts.noise || 0;   // extension: Gaussian velocity noise, world units / s
    const data = native().packParticles(pts.subarray(0, 3 * n), Float64Array.from(this.origin), this.dx, Int32Array.from(this.dims),
      this.bound, mat, Float64Array.from(vel), s, noise, dflt(opts.noiseSeed, Math.imul(this.count + 1, 2654435761 | 0) >>> 0));
    this.P.write(data, this.count * 8);
    this.count += n;
    call(() => native().setCount(this._h, this.count));
    return n;
  }

  clear() {                                                                       // mpm.js:238-242
    this.count = 0;
    if (!this.destroyed) call(() => native().setCount(this._h, 0));
    this.materials = [];
    this.materialIndex.clear();
  }

  setColliders(list) { this.colliders = list; }                                  // mpm.js:274

  colliderTable() {                                                               // mpm.js:276-293
    const data = new Float32Array(16 * 12);
    const u = new Uint32Array(data.buffer);
    const s = this.substepDt / this.dx;
    this.colliders.slice(0, 16).forEach((c, k) => {
      const o = k * 12;
      u[o] = { sphere: 0, box: 1, plane: 2 }[c.type || 'sphere'];
      const ctr = c.center || [0, 0, 0];
      for (let a = 0; a < 3; a++) data[o + 1 + a] = (ctr[a] - this.origin[a]) / this.dx;
      if (c.type === 'plane') { const nrm = c.normal || [0, 1, 0]; data[o + 4] = nrm[0]; data[o + 5] = nrm[1]; data[o + 6] = nrm[2]; }
      else if (c.type === 'box') { const h = c.half || [0.1, 0.1, 0.1]; data[o + 4] = h[0] / this.dx; data[o + 5] = h[1] / this.dx; data[o + 6] = h[2] / this.dx; }
      else data[o + 4] = dflt(c.radius, 0.1) / this.dx;
      data[o + 7] = dflt(c.friction, 0.3);
      const v = c.velocity || [0, 0, 0];
      data[o + 8] = v[0] * s; data[o + 9] = v[1] * s; data[o + 10] = v[2] * s;
    });
    return data;
  }

  latticeGravity() {                                                              // mpm.js:588
    const dt = this.substepDt;
    return Float32Array.from(this.gravity.map((x) => (x * dt * dt) / this.dx));
  }

  // ---- stepping (mpm.js:582-621) ----
  step() {
    if (this.destroyed) throw new SIError('MPM.step: it was destroyed', 'destroyed');
    if (this.count === 0) return;
    call(() => native().setColliders(this._h, this.colliderTable(), Math.min(16, this.colliders.length)));
    call(() => native().step(this._h, this.substeps, this.latticeGravity(), 0.4, 0.9));
  }

  /** MPM.sortEvery (mpm.js:101, 591): sort every n substeps, from the next substep on. */
  get sortEvery() { return this._sortEvery; }
  set sortEvery(n) {
    call(() => native().setSortEvery(this._alive('MPM.sortEvery'), n));
    this._sortEvery = n;
  }

  /** Clear the grid and scatter the particles (P2G, mpm.js:607-612); the grid is left in i32 fixed
   * point. grid: this simulation's (default) or a gpu.array('vec4i', numNodes), which receives it. */
  p2g(P, forceGlobal, grid) {
    if (P && P !== this.P) throw new SIError('MPM.p2g: only the simulation\'s own particle buffer is supported', 'unsupported');
    if (grid && grid !== this.grid) {
      if (grid.type !== 'vec4i' || grid.length !== this.numNodes)
        throw new SIError(`MPM.p2g: grid must be a vec4i array of ${this.numNodes} nodes`, 'invalid-argument');
      call(() => native().p2gTo(this._alive('MPM.p2g'), !!forceGlobal, grid.data));
      return;
    }
    call(() => native().p2g(this._alive('MPM.p2g'), !!forceGlobal));
  }

  _sort() { call(() => native().sort(this._alive('MPM._sort'))); }

  async readPositions() {                                                         // mpm.js:624-630
    // x * dx + origin is evaluated on the device in f64 and rounded to f32, as here in JS
    const pts = (await this.readPoints())[0];
    const out = new Float32Array(this.count * 3);
    for (let k = 0; k < this.count; k++)
      for (let a = 0; a < 3; a++) out[3 * k + a] = pts[4 * k + a];
    return out;
  }

  /** [points, velocities]: (world xyz, material id bits) per particle and, if asked, (world v, J). */
  readPoints(withVelocity) {
    return callAsync(() => native().readPoints(this._alive('MPM.readPoints'), this.count, Float64Array.from(this.origin),
      this.dx, this.dx / this.substepDt, !!withVelocity));
  }

  /** One substep split at the coupling call site (mpm.js:596): sort when due, P2G, grid update. */
  substepBegin() {
    call(() => native().setColliders(this._alive('MPM.substepBegin'), this.colliderTable(), Math.min(16, this.colliders.length)));
    call(() => native().substepBegin(this._h, this.latticeGravity(), 0.4));
  }

  /** G2P: ends the substep begun with substepBegin. */
  substepEnd() { call(() => native().substepEnd(this._alive('MPM.substepEnd'), 0.9)); }

  palette() { return this.materials.map((m) => m.color || [0.7, 0.7, 0.7]); }
  fluidMask() { return this.materials.reduce((m, d, k) => (d.kind === FLUID ? m | (1 << k) : m), 0); }

  destroy() {                                                                     // mpm.js:262-267
    if (this.destroyed) return;
    this._lastStepCount = native().getStepCount(this._h);
    this.destroyed = true;
    native().simDestroy(this._h);
  }
}

// sampleShape (mpm.js:683-714): box / sphere / cylinder through the library's sampler; sdf shapes with
// a JavaScript callback are sampled here with the same xorshift32 stream.
function sampleShape(shape, spacing, jitter, seed) {
  const prm = new Float64Array(6);
  let kind;
  if (shape.box) { kind = 0; prm.set(shape.box[0], 0); prm.set(shape.box[1], 3); }
  else if (shape.sphere) { kind = 1; prm.set(shape.sphere[0], 0); prm[3] = shape.sphere[1]; }
  else if (shape.cylinder) { kind = 2; prm.set(shape.cylinder[0], 0); prm[3] = shape.cylinder[1]; prm[4] = shape.cylinder[2]; }
  else if (shape.sdf) return sampleSdf(shape, spacing, jitter, seed);
  else throw new SIError('MPM.add: shape needs box | sphere | cylinder | sdf | points', 'invalid-argument');
  return call(() => native().sampleShape(kind, prm, spacing, jitter, seed >>> 0));
}

function sampleSdf(shape, spacing, jitter, seed) {
  let rnd = seed >>> 0 || 1;
  const rand = () => { rnd ^= rnd << 13; rnd ^= rnd >>> 17; rnd ^= rnd << 5; return ((rnd >>> 0) / 4294967296); };
  const lo = shape.bounds[0], hi = shape.bounds[1];
  const n = [0, 1, 2].map((a) => Math.max(1, Math.floor((hi[a] - lo[a]) / spacing)));
  const out = [];
  const p = [0, 0, 0];
  for (let z = 0; z < n[2]; z++)
    for (let y = 0; y < n[1]; y++)
      for (let x = 0; x < n[0]; x++) {
        p[0] = lo[0] + (x + 0.5 + (rand() - 0.5) * jitter) * spacing;
        p[1] = lo[1] + (y + 0.5 + (rand() - 0.5) * jitter) * spacing;
        p[2] = lo[2] + (z + 0.5 + (rand() - 0.5) * jitter) * spacing;
        if (shape.sdf(p) <= 0) out.push(p[0], p[1], p[2]);
      }
  return new Float32Array(out);
}

/** Grid mass and momentum after P2G equal the particles' (src/verify/index.js:175-193). */
async function checkMPM(sim, tol) {
  tol = tol === undefined ? 1e-4 : tol;
  const n = sim.count, raw = await sim.P.read(n * 8);
  sim.p2g();
  const g = await sim.grid.read();
  const FP = Math.pow(2, 22);
  const vol = 1 / Math.pow(sim.ppcAxis, 3);
  const u = new Uint32Array(raw.buffer, raw.byteOffset, raw.length);
  let pm = 0, px = 0, py = 0, pz = 0, gm = 0, gx = 0, gy = 0, gz = 0;
  for (let k = 0; k < n; k++) {
    const mat = sim.materials[u[32 * k + 3]];
    const m = (mat && mat.density !== undefined ? mat.density : 1) * vol;
    pm += m; px += m * raw[32 * k + 4]; py += m * raw[32 * k + 5]; pz += m * raw[32 * k + 6];
  }
  for (let k = 0; k < g.length; k += 4) { gx += g[k]; gy += g[k + 1]; gz += g[k + 2]; gm += g[k + 3]; }
  gm /= FP; gx /= FP; gy /= FP; gz /= FP;
  const massErr = Math.abs(gm - pm) / pm;
  const momErr = Math.hypot(gx - px, gy - py, gz - pz) / Math.max(1e-9, pm);
  return { ok: massErr < tol && momErr < tol + (n * 27 * 4) / FP / pm, massError: massErr, momentumError: momErr };
}

module.exports = { init, Context, MPM, MATERIALS, MPM_KINDS, SIError, checkMPM, sampleShape, PARTICLE_FLOATS, native };
