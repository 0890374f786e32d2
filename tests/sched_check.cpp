// CPU check of the serving scheduler (shaderintel_amd/csrc/serve_sched.hpp) against a fake pool with
// a small particle budget: FIFO admission without overtaking, retirement after the lifetime, rejection
// of requests that cannot fit an empty pool, ticket states, result counts, latencies stamped when a
// step's device work ends, censored ages, and the scene-step requests' latencies and ages. Exit
// status 0 = pass.
#include <cstdio>
#include <map>
#include <vector>

#include "serve_sched.hpp"

using namespace mpm_serve;

#define CHECK(c)                                                        \
  do {                                                                  \
    if (!(c)) {                                                         \
      std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
      return 1;                                                         \
    }                                                                   \
  } while (0)

struct FakePool {
  uint32_t budget, used = 0, next_scene = 0;
  std::map<uint32_t, uint32_t> size_of_tmpl, scene_size;
  std::vector<uint32_t> admitted_tmpl;
  int try_admit(const Request& r, bool empty, uint32_t* scene) {
    const uint32_t n = size_of_tmpl[r.tmpl];
    if (used + n > budget) return empty ? kNeverFits : kNoRoom;
    used += n;
    *scene = next_scene++;
    scene_size[*scene] = n;
    admitted_tmpl.push_back(r.tmpl);
    return kAdmitted;
  }
  void remove(uint32_t scene) { used -= scene_size[scene]; scene_size.erase(scene); }
};

int main() {
  FakePool pool{100};
  pool.size_of_tmpl = {{0, 60}, {1, 50}, {2, 30}, {3, 200}, {4, 10}};
  Scheduler s;
  auto admit = [&](double t) {
    return s.admit(t, [&](const Request& r, bool empty, uint32_t* sc) { return pool.try_admit(r, empty, sc); });
  };
  uint64_t step_index = 0;
  auto step = [&](double t) {  // enqueue a step and see it finish at t
    const uint64_t k = step_index++;
    const int n = s.after_step(k, t - 0.5, [&](const Request& r) {
      pool.remove(r.scene);
      return r.result ? 7u : 0u;  // records a real loop would copy
    });
    s.finish_step(k, t);
    return n;
  };

  float out[8];
  s.submit(0, 2, 0.0);             // 60 particles, 2 steps
  s.submit(1, 1, 0.1, out, 8);     // 50: does not fit next to the 60 (results wanted)
  s.submit(2, 1, 0.2);             // 30 would fit, but must not overtake
  CHECK(admit(1.0) == 1);
  CHECK(s.running() == 1 && s.waiting() == 2 && pool.used == 60);
  CHECK(s.state(0) == kRunning && s.state(1) == kWaiting && s.state(2) == kWaiting && !s.known(9));
  {  // censored ages at t = 1.5: the running one and both waiting ones by their age so far
    const std::vector<double> a = s.ages(1.5);
    CHECK(a.size() == 3 && a[0] == 1.5 && a[1] > 1.39 && a[1] < 1.41 && a[2] > 1.29 && a[2] < 1.31);
  }
  CHECK(step(2.0) == 0);
  CHECK(admit(2.0) == 0);  // still blocked behind the 50
  CHECK(step(3.0) == 1);   // the 60 retires after its 2 steps
  CHECK(pool.used == 0);
  CHECK(admit(3.0) == 2);  // 50 + 30 now
  CHECK(pool.admitted_tmpl.size() == 3 && pool.admitted_tmpl[1] == 1 && pool.admitted_tmpl[2] == 2);
  CHECK(step(4.0) == 2);
  CHECK(s.running() == 0 && s.waiting() == 0 && pool.used == 0);
  CHECK(s.state(1) == kDone && s.count(1) == 7 && s.count(2) == 0);
  // a request larger than an empty pool is rejected; the one behind it proceeds
  s.submit(3, 5, 5.0);
  s.submit(4, 1, 5.0);
  CHECK(admit(6.0) == 1);
  CHECK(s.rejected().size() == 1 && s.rejected()[0] == 3 && s.state(3) == kRejected);
  CHECK(step(7.0) == 1);
  // latency bookkeeping: ids 0..2 and 4 completed
  const auto& c = s.completed();
  CHECK(c.size() == 4);
  CHECK(c[0].id == 0 && c[0].wait_s == 1.0 && c[0].total_s == 3.0 && c[0].steps == 2 && c[0].stamped);
  CHECK(c[1].id == 1 && c[1].wait_s > 2.89 && c[1].wait_s < 2.91 && c[1].total_s > 3.89 && c[1].total_s < 3.91);
  CHECK(c[3].id == 4 && c[3].wait_s == 1.0 && c[3].total_s == 2.0);
  CHECK(s.submitted() == 5);
  // a full pool admits as soon as room appears, in order, many at once
  for (int k = 0; k < 10; k++) s.submit(4, 1 + k % 3, 10.0);  // ten scenes of 10
  CHECK(admit(10.0) == 10 && pool.used == 100);
  s.submit(2, 1, 10.5);
  CHECK(admit(11.0) == 0);
  CHECK(step(12.0) == 4);  // lifetimes 1 retire: k = 0, 3, 6, 9
  CHECK(admit(12.0) == 1 && pool.used == 90);
  {  // a retired session's latency is provisional until its step is finished
    const uint64_t k = step_index++;
    CHECK(s.after_step(k, 13.0, [&](const Request& r) { pool.remove(r.scene); return 0u; }) >= 1);
    CHECK(!s.completed().back().stamped);
    s.finish_step(k, 14.0);
    CHECK(s.completed().back().stamped && s.completed().back().total_s == 14.0 - 10.5);  // the one submitted at 10.5
  }
  {  // scene-step requests: a session's first on arrival, the next ones at the end of the previous step
    Scheduler q;
    FakePool qp{100};
    qp.size_of_tmpl = {{0, 10}};
    auto qadmit = [&](double t) {
      return q.admit(t, [&](const Request& r, bool empty, uint32_t* sc) { return qp.try_admit(r, empty, sc); });
    };
    auto qstep = [&](uint64_t k, double t_end) {
      std::vector<double> firsts;
      const uint32_t n = q.begin_step(k, &firsts);
      q.after_step(k, t_end - 0.1, [&](const Request& r) { qp.remove(r.scene); return 0u; });
      q.finish_step(k, t_end, n, &firsts);
      return n;
    };
    q.submit(0, 3, 0.0);  // A: three steps
    q.submit(0, 1, 0.2);  // B: one step
    CHECK(qadmit(0.5) == 2);
    CHECK(qstep(0, 1.0) == 2);  // A's and B's first requests: issued at 0.0 and 0.2
    q.submit(0, 1, 1.1);        // C
    CHECK(qadmit(1.2) == 1);
    CHECK(qstep(1, 1.5) == 2);  // A's second (issued at 1.0) and C's first (issued at 1.1)
    const std::vector<double>& l = q.scene_step_latencies();
    CHECK(l.size() == 4 && l[0] == 1.0 && l[1] > 0.79 && l[1] < 0.81 && l[2] > 0.39 && l[2] < 0.41 && l[3] == 0.5);
    CHECK(q.scene_steps() == 4);
    q.submit(0, 2, 1.6);  // D: waiting (not admitted yet)
    const std::vector<double> a = q.scene_step_ages(1.7);
    // the four served, then A (pending since 1.5), then D (since 1.6); C retired
    CHECK(a.size() == 6 && a[4] > 0.19 && a[4] < 0.21 && a[5] > 0.09 && a[5] < 0.11);
    CHECK(qadmit(1.7) == 1);
    const std::vector<double> b = q.scene_step_ages(1.8);  // D resident, its first request pending since 1.6
    CHECK(b.size() == 6 && b[5] > 0.19 && b[5] < 0.21);
  }
  std::printf("ok\n");
  return 0;
}
