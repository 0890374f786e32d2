// Node N-API addon: exposes the C ABI of libmpm_hip.so (include/mpm.h) to the JavaScript host
// (index.js), which keeps ShaderIntel's MPM world API (src/physics/mpm.js:82-680).
//
// Handles are napi_external values with finalizers, so a simulation that JavaScript drops is freed
// like the reference's destroy() (mpm.js:262-267). Errors are thrown as JS Errors carrying the
// reference's SIError code strings in `.code` (src/core/errors.js:3-11).
#include <node_api.h>

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mpm.h"

namespace {

struct CtxBox {
  mpm_ctx* ctx = nullptr;
  int live_sims = 0;
  bool released = false;  // JavaScript dropped or destroyed the context
};

struct SimBox {
  mpm_sim* sim = nullptr;
  CtxBox* owner = nullptr;
};

void maybe_free_ctx(CtxBox* c) {
  if (c->released && c->live_sims == 0) {
    mpm_ctx_destroy(c->ctx);
    delete c;
  }
}

void destroy_sim(SimBox* b) {
  if (b->sim) {
    mpm_sim_destroy(b->sim);
    b->sim = nullptr;
    b->owner->live_sims--;
    maybe_free_ctx(b->owner);
  }
}

void throw_status(napi_env env, mpm_status st, const char* what) {
  std::string msg = std::string("ShaderIntel: ") + what + ": " + mpm_last_error_message();
  napi_value m, code, err;
  napi_create_string_utf8(env, msg.c_str(), NAPI_AUTO_LENGTH, &m);
  napi_create_string_utf8(env, mpm_status_string(st), NAPI_AUTO_LENGTH, &code);
  napi_create_error(env, code, m, &err);
  napi_set_named_property(env, err, "code", code);
  napi_throw(env, err);
}

void throw_arg(napi_env env, const char* msg) {
  napi_value m, code, err;
  std::string s = std::string("ShaderIntel: ") + msg;
  napi_create_string_utf8(env, s.c_str(), NAPI_AUTO_LENGTH, &m);
  napi_create_string_utf8(env, "invalid-argument", NAPI_AUTO_LENGTH, &code);
  napi_create_error(env, code, m, &err);
  napi_set_named_property(env, err, "code", code);
  napi_throw(env, err);
}

#define ARGS(N)                                          \
  size_t argc = N;                                       \
  napi_value argv[N];                                    \
  napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr)

double num(napi_env env, napi_value v, double dflt = 0) {
  double d = dflt;
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t == napi_number) napi_get_value_double(env, v, &d);
  return d;
}

napi_value make_num(napi_env env, double d) {
  napi_value r;
  napi_create_double(env, d, &r);
  return r;
}

napi_value undef(napi_env env) {
  napi_value r;
  napi_get_undefined(env, &r);
  return r;
}

// typed-array access: returns data pointer and element count, or nullptr if the type differs
void* typed(napi_env env, napi_value v, napi_typedarray_type want, size_t* len) {
  bool is = false;
  napi_is_typedarray(env, v, &is);
  if (!is) return nullptr;
  napi_typedarray_type t;
  size_t n;
  void* data;
  napi_value ab;
  size_t off;
  napi_get_typedarray_info(env, v, &t, &n, &data, &ab, &off);
  if (t != want) return nullptr;
  *len = n;
  return data;
}

napi_value make_typed(napi_env env, napi_typedarray_type t, size_t n, size_t elem, void** data) {
  napi_value ab, arr;
  napi_create_arraybuffer(env, n * elem, data, &ab);
  napi_create_typedarray(env, t, n, ab, 0, &arr);
  return arr;
}

SimBox* sim_of(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    throw_arg(env, "expected a simulation handle");
    return nullptr;
  }
  SimBox* b = static_cast<SimBox*>(p);
  if (!b->sim) {
    napi_value m, code, err;
    napi_create_string_utf8(env, "ShaderIntel: MPM: it was destroyed", NAPI_AUTO_LENGTH, &m);
    napi_create_string_utf8(env, "destroyed", NAPI_AUTO_LENGTH, &code);
    napi_create_error(env, code, m, &err);
    napi_set_named_property(env, err, "code", code);
    napi_throw(env, err);
    return nullptr;
  }
  return b;
}

CtxBox* ctx_of(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    throw_arg(env, "expected a context handle");
    return nullptr;
  }
  CtxBox* c = static_cast<CtxBox*>(p);
  if (c->released) {
    throw_arg(env, "context was destroyed");
    return nullptr;
  }
  return c;
}

// A read that resolves later (the reference's P.read / grid.read, src/core/context.js:253-270): the
// snapshot is enqueued on the simulation's stream by the caller (mpm_read_*_begin, on the JS thread,
// so it is ordered with the calls around it); the wait and the copy into the result's buffer run on
// a worker thread (napi_async_work), and the Promise settles on the JS thread.
struct ReadJob {
  napi_async_work work = nullptr;
  napi_deferred deferred = nullptr;
  napi_ref result = nullptr;  // the value the Promise resolves with (holds the buffer being filled)
  mpm_read* read = nullptr;
  void* dst = nullptr;
  mpm_status st = MPM_OK;
  std::string err, what;
};

void ReadExecute(napi_env, void* data) {
  ReadJob* j = static_cast<ReadJob*>(data);
  j->st = mpm_read_finish(j->read, j->dst);
  j->read = nullptr;
  if (j->st) j->err = mpm_last_error_message();
}

void ReadComplete(napi_env env, napi_status status, void* data) {
  ReadJob* j = static_cast<ReadJob*>(data);
  napi_value v;
  napi_get_reference_value(env, j->result, &v);
  if (status == napi_ok && j->st == MPM_OK) {
    napi_resolve_deferred(env, j->deferred, v);
  } else {
    const mpm_status st = j->st ? j->st : MPM_ERR_ERROR;
    std::string msg = "ShaderIntel: " + j->what + ": " + (j->err.empty() ? "read cancelled" : j->err);
    napi_value m, code, err;
    napi_create_string_utf8(env, msg.c_str(), NAPI_AUTO_LENGTH, &m);
    napi_create_string_utf8(env, mpm_status_string(st), NAPI_AUTO_LENGTH, &code);
    napi_create_error(env, code, m, &err);
    napi_set_named_property(env, err, "code", code);
    napi_reject_deferred(env, j->deferred, err);
  }
  napi_delete_reference(env, j->result);
  napi_delete_async_work(env, j->work);
  delete j;
}

napi_value queue_read(napi_env env, mpm_read* read, void* dst, napi_value result, const char* what) {
  ReadJob* j = new ReadJob;
  j->read = read;
  j->dst = dst;
  j->what = what;
  napi_value promise, name;
  napi_create_promise(env, &j->deferred, &promise);
  napi_create_reference(env, result, 1, &j->result);
  napi_create_string_utf8(env, what, NAPI_AUTO_LENGTH, &name);
  napi_create_async_work(env, nullptr, name, ReadExecute, ReadComplete, j, &j->work);
  napi_queue_async_work(env, j->work);
  return promise;
}

// ---- context ----
napi_value CtxCreate(napi_env env, napi_callback_info info) {
  ARGS(1);
  int dev = argc > 0 ? (int)num(env, argv[0], 0) : 0;
  mpm_ctx* c = nullptr;
  mpm_status st = mpm_ctx_create(dev, &c);
  if (st) { throw_status(env, st, "init"); return nullptr; }
  CtxBox* box = new CtxBox;
  box->ctx = c;
  napi_value ext;
  napi_create_external(env, box, [](napi_env, void* data, void*) {
    CtxBox* b = static_cast<CtxBox*>(data);
    if (!b->released) { b->released = true; maybe_free_ctx(b); }
  }, nullptr, &ext);
  return ext;
}

napi_value CtxDestroy(napi_env env, napi_callback_info info) {
  ARGS(1);
  void* p = nullptr;
  if (napi_get_value_external(env, argv[0], &p) == napi_ok && p) {
    CtxBox* b = static_cast<CtxBox*>(p);
    if (!b->released) { b->released = true; maybe_free_ctx(b); }
  }
  return undef(env);
}

napi_value CtxInfo(napi_env env, napi_callback_info info) {
  ARGS(1);
  CtxBox* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  char name[256];
  int32_t cus = 0;
  mpm_ctx_info(c->ctx, name, sizeof name, &cus);
  napi_value o, s;
  napi_create_object(env, &o);
  napi_create_string_utf8(env, name, NAPI_AUTO_LENGTH, &s);
  napi_set_named_property(env, o, "name", s);
  napi_set_named_property(env, o, "cus", make_num(env, cus));
  return o;
}

// ---- simulation ----
napi_value SimCreate(napi_env env, napi_callback_info info) {
  ARGS(2);
  CtxBox* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  mpm_config cfg;
  std::memset(&cfg, 0, sizeof cfg);
  napi_value dims, v;
  napi_get_named_property(env, argv[1], "dims", &dims);
  for (uint32_t a = 0; a < 3; a++) {
    napi_get_element(env, dims, a, &v);
    cfg.dims[a] = (int32_t)num(env, v, 0);
  }
  napi_get_named_property(env, argv[1], "maxParticles", &v);
  cfg.max_particles = (uint32_t)num(env, v, 0);
  napi_get_named_property(env, argv[1], "ppcAxis", &v);
  cfg.ppc_axis = (int32_t)num(env, v, 2);
  napi_get_named_property(env, argv[1], "bound", &v);
  cfg.bound = (float)num(env, v, 3);
  napi_get_named_property(env, argv[1], "sortEvery", &v);
  cfg.sort_every = (uint32_t)num(env, v, 4);
  napi_get_named_property(env, argv[1], "flags", &v);
  cfg.flags = (uint32_t)num(env, v, 0);
  mpm_sim* s = nullptr;
  mpm_status st = mpm_sim_create(c->ctx, &cfg, &s);
  if (st) { throw_status(env, st, "MPM"); return nullptr; }
  SimBox* b = new SimBox;
  b->sim = s;
  b->owner = c;
  c->live_sims++;
  napi_value ext;
  napi_create_external(env, b, [](napi_env, void* data, void*) {
    SimBox* bb = static_cast<SimBox*>(data);
    destroy_sim(bb);
    delete bb;
  }, nullptr, &ext);
  return ext;
}

napi_value SimDestroy(napi_env env, napi_callback_info info) {
  ARGS(1);
  void* p = nullptr;
  if (napi_get_value_external(env, argv[0], &p) == napi_ok && p) destroy_sim(static_cast<SimBox*>(p));
  return undef(env);
}

napi_value SetTable(napi_env env, napi_callback_info info, bool materials) {
  ARGS(3);
  SimBox* b = sim_of(env, argv[0]);
  if (!b) return nullptr;
  size_t len = 0;
  float* data = static_cast<float*>(typed(env, argv[1], napi_float32_array, &len));
  uint32_t n = (uint32_t)num(env, argv[2], 0);
  if (!data || len < 12 * (size_t)n) { throw_arg(env, "table must be a Float32Array of 12 floats per row"); return nullptr; }
  mpm_status st = materials ? mpm_set_materials(b->sim, data, n) : mpm_set_colliders(b->sim, data, n);
  if (st) { throw_status(env, st, materials ? "MPM materials" : "MPM colliders"); return nullptr; }
  return undef(env);
}
napi_value SetMaterials(napi_env env, napi_callback_info info) { return SetTable(env, info, true); }
napi_value SetColliders(napi_env env, napi_callback_info info) { return SetTable(env, info, false); }

napi_value WriteParticles(napi_env env, napi_callback_info info) {
  ARGS(3);
  SimBox* b = sim_of(env, argv[0]);
  if (!b) return nullptr;
  size_t len = 0;
  float* data = static_cast<float*>(typed(env, argv[1], napi_float32_array, &len));
  if (!data || len % 32) { throw_arg(env, "P.write: expected a Float32Array of 32 floats per particle"); return nullptr; }
  uint32_t offset = (uint32_t)num(env, argv[2], 0);
  mpm_status st = mpm_write_particles(b->sim, data, offset, (uint32_t)(len / 32));
  if (st) { throw_status(env, st, "MPM.P.write"); return nullptr; }
  return undef(env);
}

napi_value ReadParticles(napi_env env, napi_callback_info info) {
  ARGS(3);
  SimBox* b = sim_of(env, argv[0]);
  if (!b) return nullptr;
  uint32_t offset = (uint32_t)num(env, argv[1], 0), n = (uint32_t)num(env, argv[2], 0);
  mpm_read* r = nullptr;
  mpm_status st = mpm_read_particles_begin(b->sim, offset, n, &r);
  if (st) { throw_status(env, st, "MPM.P.read"); return nullptr; }
  void* data = nullptr;
  napi_value arr = make_typed(env, napi_float32_array, (size_t)n * 32, 4, &data);
  return queue_read(env, r, data, arr, "MPM.P.read");
}

napi_value ReadGrid(napi_env env, napi_callback_info info) {
  ARGS(2);
  SimBox* b = sim_of(env, argv[0]);
  if (!b) return nullptr;
  double nodes = num(env, argv[1], 0);
  mpm_read* r = nullptr;
  mpm_status st = mpm_read_grid_begin(b->sim, &r);
  if (st) { throw_status(env, st, "MPM.grid.read"); return nullptr; }
  void* data = nullptr;
  napi_value arr = make_typed(env, napi_int32_array, (size_t)nodes * 4, 4, &data);
  return queue_read(env, r, data, arr, "MPM.grid.read");
}

// p2gTo(sim, forceGlobal, Int32Array(numNodes * 4)): P2G into another grid array (blocking)
napi_value P2GTo(napi_env env, napi_callback_info info) {
  ARGS(3);
  SimBox* b = sim_of(env, argv[0]);
  if (!b) return nullptr;
  bool fg = false;
  napi_get_value_bool(env, argv[1], &fg);
  size_t len = 0;
  int32_t* w = static_cast<int32_t*>(typed(env, argv[2], napi_int32_array, &len));
  if (!w) { throw_arg(env, "MPM.p2g: grid must be a vec4i array"); return nullptr; }
  mpm_status st = mpm_p2g_to(b->sim, fg ? 1 : 0, w);
  if (st) { throw_status(env, st, "MPM.p2g"); return nullptr; }
  return undef(env);
}

napi_value SetSortEvery(napi_env env, napi_callback_info info) {
  ARGS(2);
  SimBox* b = sim_of(env, argv[0]);
  if (!b) return nullptr;
  mpm_status st = mpm_set_sort_every(b->sim, (uint32_t)num(env, argv[1], 4));
  if (st) { throw_status(env, st, "MPM.sortEvery"); return nullptr; }
  return undef(env);
}

// readPoints(sim, n, origin Float64Array(3), dx, vscale, withVelocity) -> Promise<[Float32Array, Float32Array|null]>
napi_value ReadPoints(napi_env env, napi_callback_info info) {
  ARGS(6);
  SimBox* b = sim_of(env, argv[0]);
  if (!b) return nullptr;
  const uint32_t n = (uint32_t)num(env, argv[1], 0);
  size_t olen = 0;
  double* origin = static_cast<double*>(typed(env, argv[2], napi_float64_array, &olen));
  if (!origin || olen < 3) { throw_arg(env, "MPM.readPoints: origin must be a Float64Array(3)"); return nullptr; }
  bool with_vel = false;
  napi_get_value_bool(env, argv[5], &with_vel);
  mpm_read* r = nullptr;
  mpm_status st = mpm_read_points_begin(b->sim, 0, n, origin, num(env, argv[3], 1), num(env, argv[4], 1),
                                        with_vel ? 1 : 0, &r);
  if (st) { throw_status(env, st, "MPM.readPoints"); return nullptr; }
  // one buffer: the points, then (with velocities) the velocities
  void* data = nullptr;
  napi_value ab, pts, vel;
  napi_create_arraybuffer(env, (size_t)n * 16 * (with_vel ? 2 : 1), &data, &ab);
  napi_create_typedarray(env, napi_float32_array, (size_t)n * 4, ab, 0, &pts);
  if (with_vel) napi_create_typedarray(env, napi_float32_array, (size_t)n * 4, ab, (size_t)n * 16, &vel);
  else napi_get_null(env, &vel);
  napi_value out;
  napi_create_array_with_length(env, 2, &out);
  napi_set_element(env, out, 0, pts);
  napi_set_element(env, out, 1, vel);
  return queue_read(env, r, data, out, "MPM.readPoints");
}

// substepBegin(sim, gravity Float32Array(3), floorFriction)
napi_value SubstepBegin(napi_env env, napi_callback_info info) {
  ARGS(3);
  SimBox* b = sim_of(env, argv[0]);
  if (!b) return nullptr;
  size_t len = 0;
  float* g = static_cast<float*>(typed(env, argv[1], napi_float32_array, &len));
  if (!g || len < 3) { throw_arg(env, "MPM.substepBegin: gravity must be a Float32Array(3)"); return nullptr; }
  mpm_status st = mpm_substep_begin(b->sim, g, (float)num(env, argv[2], 0.4));
  if (st) { throw_status(env, st, "MPM.substepBegin"); return nullptr; }
  return undef(env);
}

napi_value SubstepEnd(napi_env env, napi_callback_info info) {
  ARGS(2);
  SimBox* b = sim_of(env, argv[0]);
  if (!b) return nullptr;
  mpm_status st = mpm_substep_end(b->sim, (float)num(env, argv[1], 0.9));
  if (st) { throw_status(env, st, "MPM.substepEnd"); return nullptr; }
  return undef(env);
}

// writeGrid(sim, Int32Array(numNodes * 4)): f32 bits of (v.xyz, mass) per node, reference layout
napi_value WriteGrid(napi_env env, napi_callback_info info) {
  ARGS(2);
  SimBox* b = sim_of(env, argv[0]);
  if (!b) return nullptr;
  size_t len = 0;
  int32_t* w = static_cast<int32_t*>(typed(env, argv[1], napi_int32_array, &len));
  if (!w) { throw_arg(env, "MPM.grid.write: words must be an Int32Array"); return nullptr; }
  mpm_status st = mpm_write_grid(b->sim, w);
  if (st) { throw_status(env, st, "MPM.grid.write"); return nullptr; }
  return undef(env);
}

napi_value SetCount(napi_env env, napi_callback_info info) {
  ARGS(2);
  SimBox* b = sim_of(env, argv[0]);
  if (!b) return nullptr;
  mpm_status st = mpm_set_count(b->sim, (uint32_t)num(env, argv[1], 0));
  if (st) { throw_status(env, st, "MPM.count"); return nullptr; }
  return undef(env);
}

napi_value SetStepCount(napi_env env, napi_callback_info info) {
  ARGS(2);
  SimBox* b = sim_of(env, argv[0]);
  if (!b) return nullptr;
  mpm_set_step_count(b->sim, (uint64_t)num(env, argv[1], 0));
  return undef(env);
}

napi_value GetStepCount(napi_env env, napi_callback_info info) {
  ARGS(1);
  SimBox* b = sim_of(env, argv[0]);
  if (!b) return nullptr;
  uint64_t n = 0;
  mpm_get_step_count(b->sim, &n);
  return make_num(env, (double)n);
}

napi_value Step(napi_env env, napi_callback_info info) {
  ARGS(5);
  SimBox* b = sim_of(env, argv[0]);
  if (!b) return nullptr;
  size_t len = 0;
  float* g = static_cast<float*>(typed(env, argv[2], napi_float32_array, &len));
  if (!g || len < 3) { throw_arg(env, "MPM.step: gravity must be a Float32Array(3)"); return nullptr; }
  mpm_status st = mpm_step(b->sim, (uint32_t)num(env, argv[1], 1), g, (float)num(env, argv[3], 0.4),
                           (float)num(env, argv[4], 0.9));
  if (st) { throw_status(env, st, "MPM.step"); return nullptr; }
  return undef(env);
}

napi_value P2G(napi_env env, napi_callback_info info) {
  ARGS(2);
  SimBox* b = sim_of(env, argv[0]);
  if (!b) return nullptr;
  bool fg = false;
  napi_get_value_bool(env, argv[1], &fg);
  mpm_status st = mpm_p2g(b->sim, fg ? 1 : 0);
  if (st) { throw_status(env, st, "MPM.p2g"); return nullptr; }
  return undef(env);
}

napi_value Sort(napi_env env, napi_callback_info info) {
  ARGS(1);
  SimBox* b = sim_of(env, argv[0]);
  if (!b) return nullptr;
  mpm_status st = mpm_sort(b->sim);
  if (st) { throw_status(env, st, "MPM._sort"); return nullptr; }
  return undef(env);
}

napi_value Rescale(napi_env env, napi_callback_info info) {
  ARGS(2);
  SimBox* b = sim_of(env, argv[0]);
  if (!b) return nullptr;
  mpm_status st = mpm_rescale(b->sim, (float)num(env, argv[1], 1));
  if (st) { throw_status(env, st, "MPM.retime"); return nullptr; }
  return undef(env);
}

napi_value Sync(napi_env env, napi_callback_info info) {
  ARGS(1);
  SimBox* b = sim_of(env, argv[0]);
  if (!b) return nullptr;
  mpm_status st = mpm_sync(b->sim);
  if (st) { throw_status(env, st, "idle"); return nullptr; }
  return undef(env);
}

// ---- primitives ----
napi_value ScanU32(napi_env env, napi_callback_info info) {
  ARGS(2);
  CtxBox* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  size_t len = 0;
  uint32_t* in = static_cast<uint32_t*>(typed(env, argv[1], napi_uint32_array, &len));
  if (!in) { throw_arg(env, "Scan.run: expected a Uint32Array"); return nullptr; }
  void* out = nullptr;
  napi_value arr = make_typed(env, napi_uint32_array, len, 4, &out);
  mpm_status st = mpm_scan_u32(c->ctx, in, static_cast<uint32_t*>(out), len);
  if (st) { throw_status(env, st, "Scan.run"); return nullptr; }
  return arr;
}

napi_value CountingSort(napi_env env, napi_callback_info info) {
  ARGS(3);
  CtxBox* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  size_t len = 0;
  uint32_t* keys = static_cast<uint32_t*>(typed(env, argv[1], napi_uint32_array, &len));
  if (!keys) { throw_arg(env, "SpatialGrid.build: expected a Uint32Array of keys"); return nullptr; }
  uint32_t nb = (uint32_t)num(env, argv[2], 0);
  void *order = nullptr, *start = nullptr;
  napi_value ao = make_typed(env, napi_uint32_array, len, 4, &order);
  napi_value as = make_typed(env, napi_uint32_array, (size_t)nb + 1, 4, &start);
  mpm_status st = mpm_counting_sort(c->ctx, keys, len, nb, static_cast<uint32_t*>(order), static_cast<uint32_t*>(start));
  if (st) { throw_status(env, st, "SpatialGrid.build"); return nullptr; }
  napi_value o;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "order", ao);
  napi_set_named_property(env, o, "start", as);
  return o;
}

// ---- host helpers ----
napi_value MaterialRow(napi_env env, napi_callback_info info) {
  ARGS(10);
  void* out = nullptr;
  napi_value arr = make_typed(env, napi_float32_array, 12, 4, &out);
  mpm_material_row((int32_t)num(env, argv[0]), num(env, argv[1]), num(env, argv[2]), num(env, argv[3]), num(env, argv[4]),
                   num(env, argv[5]), num(env, argv[6]), num(env, argv[7]), num(env, argv[8]), (int32_t)num(env, argv[9], 2),
                   static_cast<float*>(out));
  return arr;
}

napi_value SampleShape(napi_env env, napi_callback_info info) {
  ARGS(5);
  int32_t shape = (int32_t)num(env, argv[0]);
  size_t len = 0;
  double* prm = static_cast<double*>(typed(env, argv[1], napi_float64_array, &len));
  if (!prm || len < 6) { throw_arg(env, "sampleShape: params must be a Float64Array(6)"); return nullptr; }
  double spacing = num(env, argv[2]), jitter = num(env, argv[3]);
  uint32_t seed = (uint32_t)(int64_t)num(env, argv[4], 1);
  int64_t n = mpm_sample_shape(shape, prm, spacing, jitter, seed, nullptr, 0);
  if (n < 0) { throw_arg(env, "MPM.add: shape needs box | sphere | cylinder | sdf | points"); return nullptr; }
  void* out = nullptr;
  napi_value arr = make_typed(env, napi_float32_array, (size_t)n * 3, 4, &out);
  mpm_sample_shape(shape, prm, spacing, jitter, seed, static_cast<float*>(out), n);
  return arr;
}

// packParticles(points Float32Array, origin Float64Array(3), dx, dims Int32Array(3), bound, material,
//               vel Float64Array(3), s, noiseSigma, noiseSeed) -> Float32Array
napi_value PackParticles(napi_env env, napi_callback_info info) {
  ARGS(10);
  size_t np = 0, l1 = 0, l2 = 0, l3 = 0;
  float* pts = static_cast<float*>(typed(env, argv[0], napi_float32_array, &np));
  double* origin = static_cast<double*>(typed(env, argv[1], napi_float64_array, &l1));
  int32_t* dims = static_cast<int32_t*>(typed(env, argv[3], napi_int32_array, &l2));
  double* vel = static_cast<double*>(typed(env, argv[6], napi_float64_array, &l3));
  if (!pts || !origin || !dims || !vel || l1 < 3 || l2 < 3 || l3 < 3 || np % 3) {
    throw_arg(env, "packParticles: bad arguments");
    return nullptr;
  }
  const int64_t n = (int64_t)(np / 3);
  void* out = nullptr;
  napi_value arr = make_typed(env, napi_float32_array, (size_t)n * 32, 4, &out);
  mpm_pack_particles(pts, n, origin, num(env, argv[2]), dims, num(env, argv[4]), (uint32_t)num(env, argv[5]), vel,
                     num(env, argv[7]), num(env, argv[8]), (uint32_t)(int64_t)num(env, argv[9], 1),
                     static_cast<float*>(out));
  return arr;
}

napi_value Init(napi_env env, napi_value exports) {
  struct Fn { const char* name; napi_callback cb; };
  const Fn fns[] = {
      {"ctxCreate", CtxCreate}, {"ctxDestroy", CtxDestroy}, {"ctxInfo", CtxInfo}, {"simCreate", SimCreate},
      {"simDestroy", SimDestroy}, {"setMaterials", SetMaterials}, {"setColliders", SetColliders},
      {"writeParticles", WriteParticles}, {"readParticles", ReadParticles}, {"readGrid", ReadGrid},
      {"setCount", SetCount}, {"setStepCount", SetStepCount}, {"getStepCount", GetStepCount}, {"step", Step},
      {"p2g", P2G}, {"sort", Sort}, {"rescale", Rescale}, {"sync", Sync}, {"scanU32", ScanU32},
      {"countingSort", CountingSort}, {"materialRow", MaterialRow}, {"sampleShape", SampleShape},
      {"packParticles", PackParticles}, {"readPoints", ReadPoints}, {"substepBegin", SubstepBegin},
      {"substepEnd", SubstepEnd}, {"writeGrid", WriteGrid}, {"p2gTo", P2GTo}, {"setSortEvery", SetSortEvery},
  };
  for (const Fn& f : fns) {
    napi_value v;
    napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.cb, nullptr, &v);
    napi_set_named_property(env, exports, f.name, v);
  }
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
