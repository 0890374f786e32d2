// Host-side scheduling of the serving loop (no device code): a FIFO admission queue, the set of
// resident sessions with their remaining lifetimes, ticket states and latency accounting.
//
// A request asks for one scene (a template id) to be simulated for `lifetime` steps. Requests are
// admitted in arrival order while the pool has room; a request that does not fit blocks the ones
// behind it (no overtaking, so large scenes are not starved). When the step that ends a session's
// lifetime has been enqueued, the session leaves and frees its room for the next step's admissions
// (continuous batching); its latency is stamped once that step's device work is known to be done
// (finish_step). A request that cannot fit even into an empty pool is rejected.
//
// Each step of a session is one scene-step request, issued when the session's previous step is done
// (the first one when the session arrives): its latency runs from that moment to the end of the
// step that serves it. Every resident session takes part in every step, so a session's requests
// after its first are issued at the end of the previous step.
#pragma once
#include <cstdint>
#include <deque>
#include <vector>

namespace mpm_serve {

struct Request {
  uint64_t id = 0;
  uint32_t tmpl = 0;
  uint32_t lifetime = 1;     // steps
  uint32_t done = 0;         // steps enqueued
  double t_submit = 0;       // seconds, caller's clock
  double t_admit = -1;
  uint32_t scene = 0xffffffffu;
  float* result = nullptr;   // optional: the scene's records are copied here before it leaves
  uint32_t capacity = 0;     // records result holds
  uint64_t first_step = ~0ull;  // index of the session's first step
};

struct Completed {
  uint64_t id;
  double wait_s;   // submit -> admission
  double total_s;  // submit -> last step done (provisional until finish_step stamps it)
  uint32_t steps;
  uint64_t step;   // index of the step that ended the lifetime
  bool stamped;
};

enum Admit { kAdmitted = 0, kNoRoom = 1, kNeverFits = 2 };
enum State : uint8_t { kWaiting = 0, kRunning = 1, kDone = 2, kRejected = 3 };

class Scheduler {
 public:
  uint64_t submit(uint32_t tmpl, uint32_t lifetime, double now, float* result = nullptr, uint32_t capacity = 0) {
    Request r;
    r.id = next_id_++;
    r.tmpl = tmpl;
    r.lifetime = lifetime ? lifetime : 1;
    r.t_submit = now;
    r.result = result;
    r.capacity = capacity;
    waiting_.push_back(r);
    state_.push_back(kWaiting);
    t_submit_.push_back(now);
    count_.push_back(0);
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
        state_[r.id] = kRejected;
        waiting_.pop_front();
        continue;
      }
      r.scene = scene;
      r.t_admit = now;
      state_[r.id] = kRunning;
      running_.push_back(r);
      waiting_.pop_front();
      n++;
    }
    return n;
  }

  // Step `step` is about to be enqueued with every resident session: returns how many take part and
  // appends the arrival times of those for which it is the first step.
  uint32_t begin_step(uint64_t step, std::vector<double>* first_arrivals) {
    for (Request& r : running_)
      if (r.done == 0) {
        r.first_step = step;
        if (first_arrivals) first_arrivals->push_back(r.t_submit);
      }
    return (uint32_t)running_.size();
  }

  // Step `step` of every resident session has been enqueued: sessions at the end of their lifetime
  // are retired (retire(const Request&) reads their results if asked and removes their scene) and
  // recorded, their latency provisional (`now`) until finish_step. Returns how many retired.
  template <class Retire>
  int after_step(uint64_t step, double now, Retire&& retire) {
    int n = 0;
    size_t w = 0;
    for (size_t i = 0; i < running_.size(); i++) {
      Request& r = running_[i];
      r.done++;
      if (r.done >= r.lifetime) {
        count_[r.id] = retire(r);
        completed_.push_back({r.id, r.t_admit - r.t_submit, now - r.t_submit, r.done, step, false});
        state_[r.id] = kDone;
        n++;
      } else {
        running_[w++] = r;
      }
    }
    running_.resize(w);
    return n;
  }

  // Step `step` has finished on the device at time t: stamp the latency of the sessions it retired,
  // and record the latencies of its scene-step requests: `sessions` took part, of which those whose
  // arrival times are in first_arrivals took their first step (issued on arrival); the others were
  // issued at the end of the previous step (t_last).
  void finish_step(uint64_t step, double t, uint32_t sessions = 0, const std::vector<double>* first_arrivals = nullptr) {
    while (stamp_ < completed_.size() && completed_[stamp_].step <= step) {
      Completed& c = completed_[stamp_++];
      c.total_s = t - t_submit_[c.id];
      c.stamped = true;
    }
    const uint32_t firsts = first_arrivals ? (uint32_t)first_arrivals->size() : 0u;
    if (first_arrivals)
      for (double a : *first_arrivals) step_lat_.push_back(t - a);
    for (uint32_t k = firsts; k < sessions && has_last_; k++) step_lat_.push_back(t - t_last_);
    scene_steps_ += sessions;
    t_last_ = t;
    has_last_ = true;
    last_done_ = step;
    any_done_ = true;
  }

  // Latencies of the scene-step requests served so far (s).
  const std::vector<double>& scene_step_latencies() const { return step_lat_; }
  uint64_t scene_steps() const { return scene_steps_; }

  // Censored scene-step ages at time `now`: the served ones by their latency; each resident session's
  // pending request by its age (issued at the end of the last finished step once the session has
  // had one, else on arrival); each waiting session's first request by its age.
  std::vector<double> scene_step_ages(double now) const {
    std::vector<double> v = step_lat_;
    for (const auto& r : running_) {
      const bool had_one = any_done_ && r.first_step != ~0ull && r.first_step <= last_done_;
      v.push_back(now - (had_one ? t_last_ : r.t_submit));
    }
    for (const auto& r : waiting_) v.push_back(now - r.t_submit);
    return v;
  }

  // Censored request ages at time `now`: completed requests by their latency, waiting and running
  // ones by their age so far (a lower bound of their latency); rejected ones are left out.
  std::vector<double> ages(double now) const {
    std::vector<double> v;
    for (const auto& c : completed_) v.push_back(c.stamped ? c.total_s : now - t_submit_[c.id]);
    for (const auto& r : running_) v.push_back(now - r.t_submit);
    for (const auto& r : waiting_) v.push_back(now - r.t_submit);
    return v;
  }

  State state(uint64_t id) const { return id < state_.size() ? (State)state_[id] : kRejected; }
  uint32_t count(uint64_t id) const { return id < count_.size() ? count_[id] : 0u; }
  bool known(uint64_t id) const { return id < state_.size(); }
  size_t waiting() const { return waiting_.size(); }
  size_t running() const { return running_.size(); }
  const std::vector<Request>& resident() const { return running_; }
  const std::vector<Completed>& completed() const { return completed_; }
  const std::vector<uint64_t>& rejected() const { return rejected_; }
  uint64_t submitted() const { return next_id_; }

 private:
  uint64_t next_id_ = 0;
  size_t stamp_ = 0;
  std::vector<double> step_lat_;
  uint64_t scene_steps_ = 0, last_done_ = 0;
  double t_last_ = 0;
  bool has_last_ = false, any_done_ = false;
  std::deque<Request> waiting_;
  std::vector<Request> running_;
  std::vector<Completed> completed_;
  std::vector<uint64_t> rejected_;
  std::vector<uint8_t> state_;
  std::vector<double> t_submit_;
  std::vector<uint32_t> count_;
};

}  // namespace mpm_serve
