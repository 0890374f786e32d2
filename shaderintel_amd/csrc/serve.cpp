// Serving loop over a scene pool (include/mpm.h, mpm_server_*): one native thread owns the pool,
// admits queued requests in arrival order while the pool has room, advances every resident scene by
// one step (`substeps` substeps, one launch per phase for all scenes), and retires the scenes whose
// lifetime is over, between steps (continuous batching). Room is the pool's own capacity, which the
// caller sizes to an HBM budget with mpm_pool_bytes. Scenes come from device-resident templates, so
// admission is a device-to-device copy. No reference counterpart: ShaderIntel steps one MPM object per
// scene from the JS thread (src/physics/mpm.js:582-604).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mpm.h"
#include "serve_sched.hpp"

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct mpm_server {
  mpm_pool* pool = nullptr;
  mpm_server_config cfg{};
  std::thread th;
  std::mutex mu;
  std::condition_variable cv;
  bool stop = false, drain = false;
  mpm_serve::Scheduler sched;
  std::vector<double> step_ms;
  uint64_t steps = 0, particle_substeps = 0;
  uint32_t resident_particles = 0;
  mpm_status error = MPM_OK;
  std::string error_msg;
};

namespace {

void serve_loop(mpm_server* s) {
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    if (s->stop && (!s->drain || (s->sched.waiting() == 0 && s->sched.running() == 0))) break;
    // admission (pool calls happen on this thread only)
    s->sched.admit(now_s(), [&](const mpm_serve::Request& r, bool pool_empty, uint32_t* scene) {
      const mpm_status st = mpm_pool_add_scene_from_template(s->pool, r.tmpl, scene);
      if (st == MPM_OK) return (int)mpm_serve::kAdmitted;
      if (st == MPM_ERR_OUT_OF_MEMORY && !pool_empty) return (int)mpm_serve::kNoRoom;
      return (int)mpm_serve::kNeverFits;
    });
    if (s->sched.running() == 0) {
      if (s->stop) {
        if (s->sched.waiting() == 0) break;
        continue;
      }
      s->cv.wait(lk);
      continue;
    }
    uint32_t parts = 0;
    mpm_pool_info(s->pool, &parts, nullptr, nullptr);
    lk.unlock();
    const double t0 = now_s();
    mpm_status st = mpm_pool_step(s->pool, s->cfg.substeps, s->cfg.floor_friction, s->cfg.vmax);
    if (st == MPM_OK) st = mpm_pool_sync(s->pool);
    const double t1 = now_s();
    lk.lock();
    if (st != MPM_OK) {
      s->error = st;
      s->error_msg = mpm_last_error_message();
      break;
    }
    s->step_ms.push_back((t1 - t0) * 1e3);
    s->steps++;
    s->particle_substeps += (uint64_t)parts * s->cfg.substeps;
    s->sched.after_step(t1, [&](uint32_t scene) { mpm_pool_remove_scene(s->pool, scene); });
    mpm_pool_info(s->pool, &s->resident_particles, nullptr, nullptr);
  }
}

}  // namespace

extern "C" {

mpm_status mpm_server_create(mpm_pool* pool, const mpm_server_config* cfg, mpm_server** out) {
  if (!pool || !cfg || !out) return MPM_ERR_INVALID_ARGUMENT;
  if (cfg->substeps == 0) return MPM_ERR_INVALID_ARGUMENT;
  mpm_server* s = new (std::nothrow) mpm_server;
  if (!s) return MPM_ERR_OUT_OF_MEMORY;
  s->pool = pool;
  s->cfg = *cfg;
  try {
    s->th = std::thread(serve_loop, s);
  } catch (...) {
    delete s;
    return MPM_ERR_ERROR;
  }
  *out = s;
  return MPM_OK;
}

mpm_status mpm_server_submit(mpm_server* s, uint32_t tmpl, uint32_t lifetime_steps, uint64_t* ticket) {
  if (!s) return MPM_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop) return MPM_ERR_DESTROYED;
  const uint64_t id = s->sched.submit(tmpl, lifetime_steps, now_s());
  if (ticket) *ticket = id;
  s->cv.notify_one();
  return MPM_OK;
}

mpm_status mpm_server_stats(mpm_server* s, mpm_server_counters* st) {
  if (!s || !st) return MPM_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  std::memset(st, 0, sizeof *st);
  st->submitted = s->sched.submitted();
  st->completed = s->sched.completed().size();
  st->rejected = s->sched.rejected().size();
  st->waiting = s->sched.waiting();
  st->running = s->sched.running();
  st->steps = s->steps;
  st->particle_substeps = s->particle_substeps;
  st->resident_particles = s->resident_particles;
  st->error = s->error;
  return MPM_OK;
}

mpm_status mpm_server_latencies(mpm_server* s, int32_t kind, double* out, uint64_t capacity, uint64_t* n) {
  if (!s || !n) return MPM_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  std::vector<double> v;
  if (kind == MPM_LAT_STEP_MS) {
    v = s->step_ms;
  } else if (kind == MPM_LAT_REQUEST_S || kind == MPM_LAT_WAIT_S) {
    for (const auto& c : s->sched.completed()) v.push_back(kind == MPM_LAT_REQUEST_S ? c.total_s : c.wait_s);
  } else {
    return MPM_ERR_INVALID_ARGUMENT;
  }
  *n = v.size();
  if (out) std::memcpy(out, v.data(), sizeof(double) * (size_t)std::min<uint64_t>(capacity, v.size()));
  return MPM_OK;
}

mpm_status mpm_server_stop(mpm_server* s, int32_t drain) {
  if (!s) return MPM_ERR_INVALID_ARGUMENT;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->drain = drain != 0;
    s->cv.notify_one();
  }
  if (s->th.joinable()) s->th.join();
  return s->error;
}

void mpm_server_destroy(mpm_server* s) {
  if (!s) return;
  mpm_server_stop(s, 0);
  delete s;
}

}  // extern "C"
