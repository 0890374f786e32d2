// Serving loop over a scene pool (include/mpm.h, mpm_server_*): one native thread owns the pool,
// admits queued requests in arrival order while the pool has room, advances every resident scene by
// one step (`substeps` substeps, one launch per phase for all scenes), and retires the scenes whose
// lifetime is over, between steps (continuous batching). Room is the pool's own capacity, which the
// caller sizes to an HBM budget with mpm_pool_bytes. Scenes come from device-resident templates, so
// admission is a device-to-device copy. No reference counterpart: ShaderIntel steps one MPM object per
// scene from the JS thread (src/physics/mpm.js:582-604).
//
// The loop keeps one step in flight: lifetimes are counted on the host, so the retirements and
// admissions that follow step k, and step k + 1 itself, are enqueued on the pool's stream before the
// loop waits for step k; every pool call is stream-ordered, so the device never idles between steps
// for host bookkeeping. Step durations come from HIP events around each step; a request's latency
// is stamped when the step that ended its lifetime is seen done. With graph_after = n, the pool
// replays captured sort periods once the resident scenes have stayed the same for n steps (a change
// of scenes drops the captured graphs, so under churn the loop launches eagerly).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mpm.h"
#include "pool_internal.hpp"
#include "serve_sched.hpp"

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// One enqueued step: events around it on the pool's stream, its index, the particles it advanced,
// the sessions that took part and the arrival times of those for which it was the first step.
struct Flight {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  uint64_t step = 0;
  uint32_t particles = 0, sessions = 0;
  std::vector<double> firsts;
};

}  // namespace

struct mpm_server {
  mpm_pool* pool = nullptr;
  mpm_server_config cfg{};
  std::thread th;
  std::mutex mu;
  std::condition_variable cv;       // wakes the loop (submissions, stop)
  std::condition_variable done_cv;  // wakes mpm_server_wait (tickets retired or rejected)
  bool stop = false, drain = false, finished = false;
  mpm_serve::Scheduler sched;
  std::vector<double> step_ms;
  uint64_t steps = 0, enqueued = 0, particle_substeps = 0;
  uint32_t resident_particles = 0;
  mpm_status error = MPM_OK;
  std::string error_msg;
  std::vector<Flight> spare;  // events of finished flights, for reuse
};

namespace {

mpm_status new_flight(mpm_server* s, Flight* f) {
  if (!s->spare.empty()) {
    *f = std::move(s->spare.back());
    s->spare.pop_back();
    f->firsts.clear();
    return MPM_OK;
  }
  if (hipEventCreate(&f->e0) != hipSuccess) return MPM_ERR_ERROR;
  if (hipEventCreate(&f->e1) != hipSuccess) {
    hipEventDestroy(f->e0);
    return MPM_ERR_ERROR;
  }
  return MPM_OK;
}

// Waits for the oldest enqueued step (called with the lock released) and accounts for it (lock held).
mpm_status finish_oldest(mpm_server* s, std::deque<Flight>& fl, std::unique_lock<std::mutex>& lk) {
  Flight f = std::move(fl.front());
  fl.pop_front();
  lk.unlock();
  const hipError_t e = hipEventSynchronize(f.e1);
  const double t = now_s();
  float ms = 0.0f;
  if (e == hipSuccess) hipEventElapsedTime(&ms, f.e0, f.e1);
  lk.lock();
  if (e != hipSuccess) {
    s->spare.push_back(std::move(f));
    return MPM_ERR_ERROR;
  }
  s->step_ms.push_back((double)ms);
  s->steps++;
  s->particle_substeps += (uint64_t)f.particles * s->cfg.substeps;
  s->sched.finish_step(f.step, t, f.sessions, &f.firsts);
  s->spare.push_back(std::move(f));
  s->done_cv.notify_all();
  return MPM_OK;
}

void serve_loop(mpm_server* s) {
  const hipStream_t stream = mpm::pool_stream(s->pool);
  std::deque<Flight> fl;
  std::unique_lock<std::mutex> lk(s->mu);
  mpm_status st = MPM_OK;
  uint32_t stable = 0;  // steps since the resident scenes last changed
  bool graphs = false;
  mpm_pool_use_graphs(s->pool, 0);
  for (;;) {
    if (s->stop && (!s->drain || (s->sched.waiting() == 0 && s->sched.running() == 0))) break;
    // admission (pool calls happen on this thread only; all of them are ordered on the pool's stream)
    const int admitted = s->sched.admit(now_s(), [&](const mpm_serve::Request& r, bool pool_empty, uint32_t* scene) {
      const mpm_status a = mpm_pool_add_scene_from_template(s->pool, r.tmpl, scene);
      if (a == MPM_OK) return (int)mpm_serve::kAdmitted;
      if (a == MPM_ERR_OUT_OF_MEMORY && !pool_empty) return (int)mpm_serve::kNoRoom;
      return (int)mpm_serve::kNeverFits;
    });
    s->done_cv.notify_all();  // rejections
    if (s->sched.running() == 0) {
      if (!fl.empty()) {
        if ((st = finish_oldest(s, fl, lk)) != MPM_OK) break;
        continue;
      }
      if (s->stop) {
        if (s->sched.waiting() == 0) break;
        continue;
      }
      s->cv.wait(lk);
      continue;
    }
    if (admitted) stable = 0;
    // captured sort periods only while the resident scenes stay the same (a change drops them)
    const bool want = s->cfg.graph_after && stable >= s->cfg.graph_after;
    if (want != graphs) {
      if ((st = mpm_pool_use_graphs(s->pool, want ? 1 : 0)) != MPM_OK) break;
      graphs = want;
    }
    Flight f;
    if ((st = new_flight(s, &f)) != MPM_OK) break;
    f.step = s->enqueued++;
    f.sessions = s->sched.begin_step(f.step, &f.firsts);
    mpm_pool_info(s->pool, &f.particles, nullptr, nullptr);
    lk.unlock();
    hipEventRecord(f.e0, stream);
    st = mpm_pool_step(s->pool, s->cfg.substeps, s->cfg.floor_friction, s->cfg.vmax);
    hipEventRecord(f.e1, stream);
    lk.lock();
    if (st != MPM_OK) {
      s->spare.push_back(std::move(f));
      break;
    }
    fl.push_back(std::move(f));
    // the sessions this step ends leave now (their results, if asked for, are read first: that read
    // waits for the step), so the next step's admissions see their room
    const int retired = s->sched.after_step(f.step, now_s(), [&](const mpm_serve::Request& r) -> uint32_t {
      uint32_t n = 0;
      if (r.result && r.capacity) {
        const mpm_status rs = mpm_pool_read_scene(s->pool, r.scene, r.result, r.capacity, &n);
        if (rs != MPM_OK && st == MPM_OK) st = rs;
        n = std::min(n, r.capacity);
      }
      mpm_pool_remove_scene(s->pool, r.scene);
      return n;
    });
    if (st != MPM_OK) break;
    stable = retired ? 0 : stable + 1;
    mpm_pool_info(s->pool, &s->resident_particles, nullptr, nullptr);
    while (fl.size() > 1 && st == MPM_OK) st = finish_oldest(s, fl, lk);  // one step stays in flight
    if (st != MPM_OK) break;
  }
  while (!fl.empty() && st == MPM_OK) st = finish_oldest(s, fl, lk);
  if (st != MPM_OK && s->error == MPM_OK) {
    s->error = st;
    s->error_msg = mpm_last_error_message();
  }
  s->finished = true;
  s->done_cv.notify_all();
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

mpm_status mpm_server_submit_result(mpm_server* s, uint32_t tmpl, uint32_t lifetime_steps, float* result,
                                    uint32_t capacity, uint64_t* ticket) {
  if (!s || (capacity && !result)) return MPM_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop) return MPM_ERR_DESTROYED;
  const uint64_t id = s->sched.submit(tmpl, lifetime_steps, now_s(), result, capacity);
  if (ticket) *ticket = id;
  s->cv.notify_one();
  return MPM_OK;
}

mpm_status mpm_server_submit(mpm_server* s, uint32_t tmpl, uint32_t lifetime_steps, uint64_t* ticket) {
  return mpm_server_submit_result(s, tmpl, lifetime_steps, nullptr, 0, ticket);
}

mpm_status mpm_server_wait(mpm_server* s, uint64_t ticket, int32_t timeout_ms, int32_t* state, uint32_t* count) {
  if (!s || !state) return MPM_ERR_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lk(s->mu);
  if (!s->sched.known(ticket)) return MPM_ERR_INVALID_ARGUMENT;
  auto settled = [&] {
    const mpm_serve::State q = s->sched.state(ticket);
    if (q == mpm_serve::kRejected) return true;
    if (q != mpm_serve::kDone) return s->finished;
    for (const auto& c : s->sched.completed())  // done once its latency is stamped (its step finished)
      if (c.id == ticket) return c.stamped || s->finished;
    return s->finished;
  };
  if (timeout_ms < 0)
    s->done_cv.wait(lk, settled);
  else
    s->done_cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), settled);
  *state = (int32_t)s->sched.state(ticket);
  if (count) *count = s->sched.count(ticket);
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
  st->scene_steps = s->sched.scene_steps();
  return MPM_OK;
}

mpm_status mpm_server_latencies(mpm_server* s, int32_t kind, double* out, uint64_t capacity, uint64_t* n) {
  if (!s || !n) return MPM_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  std::vector<double> v;
  if (kind == MPM_LAT_STEP_MS) {
    v = s->step_ms;
  } else if (kind == MPM_LAT_REQUEST_S || kind == MPM_LAT_WAIT_S) {
    for (const auto& c : s->sched.completed())
      if (c.stamped) v.push_back(kind == MPM_LAT_REQUEST_S ? c.total_s : c.wait_s);
  } else if (kind == MPM_LAT_REQUEST_AGE_S) {
    v = s->sched.ages(now_s());
  } else if (kind == MPM_LAT_SCENE_STEP_S) {
    v = s->sched.scene_step_latencies();
  } else if (kind == MPM_LAT_SCENE_STEP_AGE_S) {
    v = s->sched.scene_step_ages(now_s());
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
  for (Flight& f : s->spare) {
    hipEventDestroy(f.e0);
    hipEventDestroy(f.e1);
  }
  delete s;
}

}  // extern "C"
