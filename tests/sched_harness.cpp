// Test harness for the job scheduler (xpfc_cuda_amd/csrc/xpfc_sched.cpp): a
// stand-in engine that keeps a step counter per slot instead of a device, so
// the scheduling logic runs on the host.  Built by tests/test_sched_host.py
// with g++; test code only.
#include <algorithm>
#include <cmath>
#include <map>
#include <vector>

#include "../xpfc_cuda_amd/csrc/xpfc_sched.hpp"

using namespace xpfc::sched;

namespace {

struct Fake : Engine {
  std::vector<std::vector<long long>> steps;   // [class][slot] steps taken by the slot's job
  std::vector<std::vector<long long>> job;     // [class][slot] id of the job in it
  std::vector<std::vector<int>> running;
  std::map<long long, long long> diverge_at;   // job id -> steps after which it is non-finite
  int busy_polls = 0, busy_left = 0;
  std::vector<long long>* log = nullptr;       // (kind, a, b, c) quads, shared by the lanes
  int lane = 0;
  double* clock = nullptr;

  int start(const Job& j, int k, int slot) override {
    steps[k][slot] = 0;
    job[k][slot] = j.id;
    log->insert(log->end(), {1, j.id, k, slot});
    return 0;
  }
  int set_running(const std::vector<std::vector<int>>& s) override {
    running = s;
    return 0;
  }
  int launch(int len, const std::vector<std::vector<int>>& budget) override {
    long long sum = 0, nrun = 0;
    for (size_t k = 0; k < running.size(); ++k)
      for (int s : running[k]) {
        if (budget[k][s] > len) return 1;  // a budget longer than its block
        steps[k][s] += budget[k][s];
        sum += budget[k][s];
        ++nrun;
      }
    log->insert(log->end(), {2 + 10 * lane, len, nrun, sum});
    busy_left = busy_polls;
    return 0;
  }
  int query() override {
    if (busy_left > 0) {
      --busy_left;
      return 0;
    }
    return 1;
  }
  int probe(std::vector<std::vector<int>>& finite) override {
    finite.assign(steps.size(), {});
    for (size_t k = 0; k < steps.size(); ++k) {
      finite[k].assign(steps[k].size(), 1);
      for (size_t s = 0; s < steps[k].size(); ++s) {
        auto it = diverge_at.find(job[k][s]);
        if (it != diverge_at.end() && steps[k][s] >= it->second) finite[k][s] = 0;
      }
    }
    return 0;
  }
  int finish(const Job& j, int k, int slot, bool diverged, Result* r) override {
    r->F_ref = (double)steps[k][slot];  // the slot's step count, to check runs
    r->mean = diverged ? NAN : 1.0;
    log->insert(log->end(), {3, j.id, k, slot});
    return 0;
  }
};

struct Harness {
  std::vector<Fake> eng;   // one per lane
  std::vector<long long> log;
  double t = 0.0;
  Scheduler* sch = nullptr;
};

}  // namespace

extern "C" {

// lane[k]: the lane of class k (nlanes engines); lane == nullptr puts every
// class in lane 0.  busy[l]: polls a block of lane l stays in flight.
void* h_create(int ncls, const int* slots, const int* lane, int nlanes, double budget, int block, int lookahead,
               const int* busy) {
  Harness* h = new Harness();
  Config c;
  c.slots.assign(slots, slots + ncls);
  c.lane.assign((size_t)ncls, 0);
  if (lane) c.lane.assign(lane, lane + ncls);
  c.budget_bytes = budget;
  c.block = block;
  c.lookahead = lookahead;
  h->eng.resize((size_t)nlanes);
  std::vector<Engine*> engines;
  for (int l = 0; l < nlanes; ++l) {
    Fake& e = h->eng[(size_t)l];
    e.steps.resize(ncls);
    e.job.resize(ncls);
    for (int k = 0; k < ncls; ++k) {
      e.steps[k].assign(slots[k], 0);
      e.job[k].assign(slots[k], -1);
    }
    e.busy_polls = busy[l];
    e.log = &h->log;
    e.lane = l;
    engines.push_back(&e);
  }
  h->sch = new Scheduler(c, engines, [h]() { return h->t; });
  return h;
}

void h_destroy(void* p) {
  Harness* h = static_cast<Harness*>(p);
  delete h->sch;
  delete h;
}

int h_submit(void* p, long long id, int cls, long long steps, double footprint, long long cells, double arrival) {
  Harness* h = static_cast<Harness*>(p);
  Job j;
  j.id = id;
  j.cls = cls;
  j.steps = steps;
  j.footprint = (size_t)footprint;
  j.cells = cells;
  j.arrival = arrival;
  return h->sch->submit(j);
}

// one poll; the clock advances by one unit per poll
int h_poll(void* p) {
  Harness* h = static_cast<Harness*>(p);
  h->t += 1.0;
  return h->sch->poll();
}

int h_idle(void* p) { return static_cast<Harness*>(p)->sch->idle() ? 1 : 0; }

void h_diverge_at(void* p, long long id, long long steps) {
  for (Fake& e : static_cast<Harness*>(p)->eng) e.diverge_at[id] = steps;
}

int h_results(void* p, long long* id, int* status, long long* done, double* F, double* t_adm, double* t_fin,
              int max) {
  Harness* h = static_cast<Harness*>(p);
  std::vector<Result> r = h->sch->take_results();
  int n = 0;
  for (const Result& x : r) {
    if (n >= max) break;
    id[n] = x.id;
    status[n] = x.status;
    done[n] = x.steps_done;
    F[n] = x.F_ref;
    t_adm[n] = x.t_admitted;
    t_fin[n] = x.t_finished;
    ++n;
  }
  return n;
}

int h_log(void* p, long long* out, int max) {
  Harness* h = static_cast<Harness*>(p);
  const int n = (int)std::min<size_t>((size_t)max, h->log.size());
  for (int i = 0; i < n; ++i) out[i] = h->log[(size_t)i];
  return (int)h->log.size();
}

double h_stat(void* p, int which) {
  const Stats& s = static_cast<Harness*>(p)->sch->stats();
  switch (which) {
    case 0: return s.peak_resident_bytes;
    case 1: return s.resident_bytes;
    case 2: return (double)s.queued;
    case 3: return (double)s.blocks;
    case 4: return s.cell_updates;
    case 5: return (double)s.rejected;
    default: return -1.0;
  }
}
}
