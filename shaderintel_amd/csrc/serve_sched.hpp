// Host-side scheduling of the serving loop (no device code): a FIFO admission queue, the set of
// resident sessions with their remaining lifetimes, and latency accounting.
//
// A request asks for one scene (a template id) to be simulated for `lifetime` steps. Requests are
// admitted in arrival order while the pool has room; a request that does not fit blocks the ones
// behind it (no overtaking, so large scenes are not starved). Between steps, sessions whose lifetime
// is over leave and free their room (continuous batching). A request that cannot fit even into an
// empty pool is rejected.
#pragma once
#include <cstdint>
#include <deque>
#include <vector>

namespace mpm_serve {

struct Request {
  uint64_t id = 0;
  uint32_t tmpl = 0;
  uint32_t lifetime = 1;   // steps
  uint32_t done = 0;       // steps completed
  double t_submit = 0;     // seconds, caller's clock
  double t_admit = -1;
  uint32_t scene = 0xffffffffu;
};

struct Completed {
  uint64_t id;
  double wait_s;   // submit -> admission
  double total_s;  // submit -> last step done
  uint32_t steps;
};

enum Admit { kAdmitted = 0, kNoRoom = 1, kNeverFits = 2 };

class Scheduler {
 public:
  uint64_t submit(uint32_t tmpl, uint32_t lifetime, double now) {
    Request r;
    r.id = next_id_++;
    r.tmpl = tmpl;
    r.lifetime = lifetime ? lifetime : 1;
    r.t_submit = now;
    waiting_.push_back(r);
    return r.id;
  }

  // try_admit(const Request&, bool pool_empty, uint32_t* scene) -> Admit. Returns how many were admitted.
  template <class TryAdmit>
  int admit(double now, TryAdmit&& try_admit) {
    int n = 0;
    while (!waiting_.empty()) {
      Request& r = waiting_.front();
      uint32_t scene = 0xffffffffu;
      const int a = try_admit(r, running_.empty(), &scene);
      if (a == kNoRoom) break;
      if (a == kNeverFits) {
        rejected_.push_back(r.id);
        waiting_.pop_front();
        continue;
      }
      r.scene = scene;
      r.t_admit = now;
      running_.push_back(r);
      waiting_.pop_front();
      n++;
    }
    return n;
  }

  // One step of every resident session is done: sessions at the end of their lifetime are removed
  // (remove(scene)) and recorded. Returns how many retired.
  template <class Remove>
  int after_step(double now, Remove&& remove) {
    int n = 0;
    size_t w = 0;
    for (size_t i = 0; i < running_.size(); i++) {
      Request& r = running_[i];
      r.done++;
      if (r.done >= r.lifetime) {
        remove(r.scene);
        completed_.push_back({r.id, r.t_admit - r.t_submit, now - r.t_submit, r.done});
        n++;
      } else {
        running_[w++] = r;
      }
    }
    running_.resize(w);
    return n;
  }

  size_t waiting() const { return waiting_.size(); }
  size_t running() const { return running_.size(); }
  const std::vector<Request>& resident() const { return running_; }
  const std::vector<Completed>& completed() const { return completed_; }
  const std::vector<uint64_t>& rejected() const { return rejected_; }
  uint64_t submitted() const { return next_id_; }

 private:
  uint64_t next_id_ = 0;
  std::deque<Request> waiting_;
  std::vector<Request> running_;
  std::vector<Completed> completed_;
  std::vector<uint64_t> rejected_;
};

}  // namespace mpm_serve
