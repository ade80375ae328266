// Job scheduler of the XPFC service (see xpfc_sched.hpp).
#include "xpfc_sched.hpp"

#include <algorithm>

#include "../../include/xpfc.h"

namespace xpfc {
namespace sched {

Scheduler::Scheduler(Config cfg, std::vector<Engine*> engines, std::function<double()> clock)
    : cfg_(std::move(cfg)), clock_(std::move(clock)) {
  if (cfg_.block < 1) cfg_.block = 1;
  if (cfg_.lookahead < 1) cfg_.lookahead = 1;
  const size_t nc = cfg_.slots.size();
  lane_of_.resize(nc);
  for (size_t k = 0; k < nc; ++k) lane_of_[k] = cfg_.lane.empty() ? (int)k : cfg_.lane[k];
  lanes_.resize(engines.size());
  for (size_t l = 0; l < engines.size(); ++l) lanes_[l].eng = engines[l];
  run_.resize(nc);
  free_.resize(nc);
  budget_.resize(nc);
  for (size_t k = 0; k < nc; ++k) {
    run_[k].resize((size_t)std::max(0, cfg_.slots[k]));
    budget_[k].assign(run_[k].size(), 0);
    for (int s = cfg_.slots[k] - 1; s >= 0; --s) free_[k].push_back(s);  // slot 0 first
    const int l = lane_of_[k];
    if (l >= 0 && l < (int)lanes_.size()) lanes_[(size_t)l].classes.push_back((int)k);
    else run_[k].clear();  // no engine: the class takes no jobs
  }
}

int Scheduler::lane_running(const Lane& l) const {
  int n = 0;
  for (int k : l.classes)
    for (const Run& r : run_[(size_t)k]) n += r.used ? 1 : 0;
  return n;
}

int Scheduler::running_count() const {
  int n = 0;
  for (const Lane& l : lanes_) n += lane_running(l);
  return n;
}

bool Scheduler::idle() const {
  for (const Lane& l : lanes_)
    if (l.in_flight) return false;
  return queue_.empty() && running_count() == 0;
}

int Scheduler::submit(const Job& job) {
  ++st_.submitted;
  const bool known = job.cls >= 0 && job.cls < (int)run_.size() && !run_[(size_t)job.cls].empty();
  if (!known || job.steps < 1 || (double)job.footprint > cfg_.budget_bytes) {
    Result r;
    r.id = job.id;
    r.status = kRejected;
    r.t_arrival = job.arrival;
    r.t_admitted = r.t_finished = clock_();
    results_.push_back(r);
    ++st_.rejected;
    return XPFC_ERR_ARG;
  }
  queue_.push_back(job);
  st_.queued = (long long)queue_.size();
  st_.queue_max = std::max(st_.queue_max, st_.queued);
  return XPFC_OK;
}

std::vector<Result> Scheduler::take_results() {
  std::vector<Result> out;
  out.swap(results_);
  return out;
}

// The lane's block has finished: advance its jobs' counters, stop the
// diverged ones, complete the finished ones.
int Scheduler::finish_block(Lane& l) {
  // the explicit nonlinear term diverges for strongly crystallising
  // parameters: such a job stops at the end of the block it diverged in
  std::vector<std::vector<int>> finite;
  if (int rc = l.eng->probe(finite)) return -rc;
  for (int k : l.classes)
    for (size_t s = 0; s < run_[(size_t)k].size(); ++s) {
      Run& r = run_[(size_t)k][s];
      if (!r.used) continue;
      r.left -= budget_[(size_t)k][s];
      r.taken += budget_[(size_t)k][s];
      if ((size_t)k < finite.size() && s < finite[(size_t)k].size() && !finite[(size_t)k][s]) {
        r.diverged = true;
        r.left = 0;
      }
      if (r.left > 0) continue;
      Result res;
      res.id = r.job.id;
      res.status = r.diverged ? kNonfinite : kDone;
      res.steps_done = r.taken;
      if (int rc = l.eng->finish(r.job, k, (int)s, r.diverged, &res)) return -rc;
      res.t_arrival = r.job.arrival;
      res.t_admitted = r.admitted;
      res.t_finished = clock_();
      if (res.status == kDone) ++st_.done;
      else ++st_.nonfinite;
      results_.push_back(res);
      st_.resident_bytes -= (double)r.job.footprint;
      r.used = false;
      free_[(size_t)k].push_back((int)s);
      l.changed = true;
    }
  return 0;
}

int Scheduler::admit() {
  int admitted = 0;
  for (auto it = queue_.begin(); it != queue_.end();) {
    const Job& j = *it;
    Lane& l = lanes_[(size_t)lane_of_[(size_t)j.cls]];
    std::vector<int>& fr = free_[(size_t)j.cls];
    if (l.in_flight || fr.empty()) {  // its lane is mid-block or its class full: later jobs may go
      ++it;
      continue;
    }
    if (st_.resident_bytes + (double)j.footprint > cfg_.budget_bytes) break;  // memory in arrival order
    const int slot = fr.back();
    if (int rc = l.eng->start(j, j.cls, slot)) return -rc;
    fr.pop_back();
    Run& r = run_[(size_t)j.cls][(size_t)slot];
    r.job = j;
    r.left = j.steps;
    r.taken = 0;
    r.admitted = clock_();
    r.diverged = false;
    r.used = true;
    st_.resident_bytes += (double)j.footprint;
    st_.peak_resident_bytes = std::max(st_.peak_resident_bytes, st_.resident_bytes);
    l.changed = true;
    ++admitted;
    it = queue_.erase(it);
  }
  st_.queued = (long long)queue_.size();
  return admitted;
}

int Scheduler::launch_block(Lane& l) {
  if (l.changed) {
    std::vector<std::vector<int>> slots(run_.size());
    for (int k : l.classes)
      for (size_t s = 0; s < run_[(size_t)k].size(); ++s)
        if (run_[(size_t)k][s].used) slots[(size_t)k].push_back((int)s);
    if (int rc = l.eng->set_running(slots)) return -rc;
    l.changed = false;
  }
  // a block of `cap` steps (lookahead blocks while no queued job waits for
  // the lane), or fewer when every job ends sooner: the shortest power of two
  // that holds the longest budget
  bool waiting = false;
  for (const Job& j : queue_) waiting = waiting || lane_of_[(size_t)j.cls] == (int)(&l - lanes_.data());
  const int cap = waiting ? cfg_.block : cfg_.block * cfg_.lookahead;
  std::vector<std::vector<int>> budget(run_.size());
  long long most = 0;
  for (int k : l.classes) {
    for (size_t s = 0; s < run_[(size_t)k].size(); ++s) {
      const Run& r = run_[(size_t)k][s];
      budget_[(size_t)k][s] = r.used ? (int)std::min<long long>(r.left, cap) : 0;
      most = std::max<long long>(most, budget_[(size_t)k][s]);
    }
    budget[(size_t)k] = budget_[(size_t)k];
  }
  int len = 1;
  while (len < most) len *= 2;
  len = std::min(len, cap);
  if (int rc = l.eng->launch(len, budget)) return -rc;
  for (int k : l.classes)
    for (size_t s = 0; s < run_[(size_t)k].size(); ++s)
      if (run_[(size_t)k][s].used)
        st_.cell_updates += (double)budget_[(size_t)k][s] * (double)run_[(size_t)k][s].job.cells;
  ++st_.blocks;
  l.in_flight = true;
  return 1;
}

int Scheduler::poll() {
  int worked = 0;
  for (Lane& l : lanes_) {
    if (!l.in_flight) continue;
    const int q = l.eng->query();
    if (q < 0) return q;
    if (q == 0) continue;
    l.in_flight = false;
    worked = 1;
    if (int rc = finish_block(l)) return rc;
  }
  const int a = admit();
  if (a < 0) return a;
  if (a > 0) worked = 1;
  for (Lane& l : lanes_) {
    if (l.in_flight || lane_running(l) == 0) continue;
    const int rc = launch_block(l);
    if (rc < 0) return rc;
    worked = 1;
  }
  st_.running = running_count();
  return worked;
}

}  // namespace sched
}  // namespace xpfc
