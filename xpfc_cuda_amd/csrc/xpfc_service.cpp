// The job service behind the C ABI: the scheduler of xpfc_sched.cpp driving
// one context per lane whose instances are the slots of the lane's grid
// classes (a batch context with the autotuned one-grid kernels for a lane of
// one class, a context of several grids otherwise), so each block of a lane
// is one captured graph with one launch per pass over its grids; lanes run on
// their own streams.  Only the public C ABI of the contexts is used here.
#include <memory>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <new>
#include <thread>
#include <vector>

#include "../../include/xpfc.h"
#include "xpfc_host.hpp"
#include "xpfc_sched.hpp"

using namespace xpfc;

namespace {

// (dim, n0, n1, n2, L0, L1, L2) of a grid, 2-D normalised
std::vector<double> class_key(int dim, const int* n, const double* L) {
  return {(double)dim, (double)n[0], (double)n[1], dim == 3 ? (double)n[2] : 1.0, L[0], L[1],
          dim == 3 ? L[2] : 1.0};
}

// Jobs' requests and kept fields, shared by the lanes.
struct JobStore {
  bool keep_fields = false;
  std::map<long long, std::vector<double>> fields;
  std::map<long long, xpfc_job> specs;  // queued / running jobs' requests
};

class LaneEngine : public sched::Engine {
 public:
  xpfc_ctx* ctx = nullptr;
  std::vector<std::vector<int>> inst;   // [class][slot] -> instance number (empty: another lane's class)
  std::vector<xpfc_params> base;        // per class: the grid's default parameters
  JobStore* store = nullptr;
  std::vector<int> steps;               // [instance] budget of the block
  std::vector<int> finite_buf;

  ~LaneEngine() override {
    if (ctx) xpfc_destroy(ctx);
  }
  int start(const sched::Job& job, int k, int slot) override {
    const xpfc_job& j = store->specs.at(job.id);
    const int s = inst[(size_t)k][(size_t)slot];
    xpfc_params p = j.params;
    p.t_sample = 0;                      // blocks carry no sampled steps
    p.precision = base[(size_t)k].precision;
    p.device = base[(size_t)k].device;
    if (int rc = xpfc_inst_set_params(ctx, s, &p)) return rc;
    if (int rc = xpfc_inst_reset(ctx, s)) return rc;
    // the admission path allocates nothing on the device or host
    return xpfc_inst_initialize_noise_noalloc(ctx, s, j.amplitude, j.seed);
  }
  int set_running(const std::vector<std::vector<int>>& slots) override {
    std::vector<int> a;
    for (size_t k = 0; k < slots.size(); ++k)
      for (int sl : slots[k]) a.push_back(inst[k][(size_t)sl]);
    const int none = 0;
    return xpfc_set_active(ctx, a.empty() ? &none : a.data(), (int)a.size());
  }
  int launch(int len, const std::vector<std::vector<int>>& budget) override {
    std::fill(steps.begin(), steps.end(), 0);
    for (size_t k = 0; k < budget.size(); ++k)
      for (size_t sl = 0; sl < budget[k].size(); ++sl) steps[(size_t)inst[k][sl]] = budget[k][sl];
    return xpfc_step_budget(ctx, len, steps.data());
  }
  int query() override { return xpfc_query(ctx); }
  int probe(std::vector<std::vector<int>>& finite) override {
    if (int rc = xpfc_probe_finite(ctx, finite_buf.data())) return rc;
    finite.resize(inst.size());
    for (size_t k = 0; k < inst.size(); ++k) {
      finite[k].resize(inst[k].size());
      for (size_t sl = 0; sl < inst[k].size(); ++sl) finite[k][sl] = finite_buf[(size_t)inst[k][sl]];
    }
    return XPFC_OK;
  }
  int finish(const sched::Job& job, int k, int slot, bool diverged, sched::Result* r) override {
    const int s = inst[(size_t)k][(size_t)slot];
    r->F_ref = r->f_mean = NAN;
    if (!diverged) {
      const int rc = xpfc_inst_fe_calc(ctx, s, &r->F_ref, &r->f_mean);
      if (rc == XPFC_ERR_NONFINITE) r->status = sched::kNonfinite;
      else if (rc) return rc;
    }
    if (int rc = xpfc_inst_avg_calc(ctx, s, &r->mean)) return rc;
    if (store->keep_fields) {
      std::vector<double>& f = store->fields[job.id];
      f.resize((size_t)job.cells);
      if (int rc = xpfc_inst_get_field(ctx, s, f.data())) return rc;
    }
    store->specs.erase(job.id);
    return XPFC_OK;
  }
};

}  // namespace

struct xpfc_service {
  JobStore store;
  std::vector<std::unique_ptr<LaneEngine>> lanes;
  std::vector<std::vector<double>> keys;  // per class
  std::vector<size_t> footprint;          // per class
  std::vector<long long> cells;           // per class
  sched::Scheduler* sch = nullptr;
  std::chrono::steady_clock::time_point t0;
  std::vector<sched::Result> ready;       // finished, not yet handed out
  double clock() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

extern "C" size_t xpfc_service_struct_size(int which) {
  switch (which) {
    case 0: return sizeof(xpfc_service_class);
    case 1: return sizeof(xpfc_service_config);
    case 2: return sizeof(xpfc_job);
    case 3: return sizeof(xpfc_job_result);
    case 4: return sizeof(xpfc_service_stats);
    default: return 0;
  }
}

extern "C" int xpfc_service_create(const xpfc_service_config* cfg, xpfc_service** out) {
  if (!cfg || !out || cfg->nclasses < 1 || !cfg->classes) return fail(XPFC_ERR_ARG, "bad service configuration");
  *out = nullptr;
  if (!(cfg->budget_bytes > 0.0)) return fail(XPFC_ERR_ARG, "budget must be positive");
  xpfc_service* v = new (std::nothrow) xpfc_service();
  if (!v) return fail(XPFC_ERR_NOMEM, "host allocation failed");
  std::vector<std::vector<xpfc_params>> ps;  // per lane: its instances
  std::vector<xpfc_params> base;
  std::vector<std::vector<std::vector<int>>> inst;  // [lane][class][slot]
  std::map<int, int> lane_no;               // the caller's lane numbers, densely
  std::vector<int> lane_of;
  sched::Config sc;
  sc.budget_bytes = cfg->budget_bytes;
  sc.block = cfg->block_steps > 0 ? cfg->block_steps : 50;
  sc.lookahead = cfg->lookahead > 0 ? cfg->lookahead : 1;
  for (int k = 0; k < cfg->nclasses; ++k) {
    const xpfc_service_class& c = cfg->classes[k];
    if (c.slots < 1) {
      delete v;
      return fail(XPFC_ERR_ARG, "class %d: slots must be >= 1", k);
    }
    xpfc_params p;
    xpfc_params_default(&p);
    p.dim = c.dim;
    for (int d = 0; d < 3; ++d) {
      p.n[d] = c.n[d];
      p.L[d] = c.L[d];
    }
    if (p.dim == 2) {
      p.n[2] = 1;
      p.L[2] = 1.0;
    }
    p.precision = cfg->precision;
    p.device = cfg->device;
    p.t_sample = 0;
    const size_t fp = xpfc_instance_bytes(&p);
    if (fp == 0) {
      delete v;
      return fail(XPFC_ERR_ARG, "class %d: unsupported grid", k);
    }
    v->keys.push_back(class_key(p.dim, p.n, p.L));
    v->footprint.push_back(fp);
    v->cells.push_back((long long)p.n[0] * p.n[1] * (p.dim == 3 ? p.n[2] : 1));
    base.push_back(p);
    sc.slots.push_back(c.slots);
    lane_of.push_back(lane_no.emplace(c.lane, (int)lane_no.size()).first->second);
  }
  const size_t nl = lane_no.size();
  ps.resize(nl);
  inst.assign(nl, std::vector<std::vector<int>>(base.size()));
  for (size_t k = 0; k < base.size(); ++k) {
    const int l = lane_of[k];
    for (int s = 0; s < sc.slots[k]; ++s) {
      inst[(size_t)l][k].push_back((int)ps[(size_t)l].size());
      ps[(size_t)l].push_back(base[k]);
    }
  }
  sc.lane = lane_of;
  for (size_t k = 0; k < v->keys.size(); ++k)
    for (size_t j = 0; j < k; ++j)
      if (v->keys[j] == v->keys[k]) {
        delete v;
        return fail(XPFC_ERR_ARG, "classes %zu and %zu have the same grid", j, k);
      }
  v->store.keep_fields = cfg->keep_fields != 0;
  std::vector<sched::Engine*> engines;
  for (size_t l = 0; l < nl; ++l) {
    std::unique_ptr<LaneEngine> e(new (std::nothrow) LaneEngine());
    if (!e) {
      delete v;
      return fail(XPFC_ERR_NOMEM, "host allocation failed");
    }
    e->inst = inst[l];
    e->base = base;
    e->store = &v->store;
    e->steps.assign(ps[l].size(), 0);
    e->finite_buf.assign(ps[l].size(), 1);
    size_t ncls = 0;
    for (const auto& k : inst[l]) ncls += k.empty() ? 0 : 1;
    const int rc0 = ncls == 1 ? xpfc_create_batch(ps[l].data(), (int)ps[l].size(), &e->ctx)
                              : xpfc_create_mixed(ps[l].data(), (int)ps[l].size(), &e->ctx);
    if (int rc = rc0) {
      delete v;
      return rc;
    }
    // every slot needs a field before the first block: the slots stay
    // inactive until a job is admitted, and admission sets the field
    const int none = 0;
    if (int rc = xpfc_set_active(e->ctx, &none, 0)) {
      delete v;
      return rc;
    }
    engines.push_back(e.get());
    v->lanes.push_back(std::move(e));
  }
  v->t0 = std::chrono::steady_clock::now();
  v->sch = new (std::nothrow) sched::Scheduler(sc, engines, [v]() { return v->clock(); });
  if (!v->sch) {
    delete v;
    return fail(XPFC_ERR_NOMEM, "host allocation failed");
  }
  *out = v;
  return XPFC_OK;
}

extern "C" void xpfc_service_destroy(xpfc_service* v) {
  if (!v) return;
  delete v->sch;
  delete v;  // the lanes' contexts with it
}

extern "C" int xpfc_service_submit(xpfc_service* v, const xpfc_job* j) {
  if (!v || !j) return fail(XPFC_ERR_ARG, "null argument");
  sched::Job job;
  job.id = j->id;
  job.steps = j->steps;
  job.arrival = j->arrival >= 0.0 ? j->arrival : v->clock();
  const std::vector<double> key = class_key(j->params.dim, j->params.n, j->params.L);
  for (size_t k = 0; k < v->keys.size(); ++k)
    if (v->keys[k] == key) {
      job.cls = (int)k;
      job.footprint = v->footprint[k];
      job.cells = v->cells[k];
    }
  if (v->store.specs.count(j->id)) return fail(XPFC_ERR_ARG, "job %lld is already queued or running", j->id);
  const bool physics_ok = std::isfinite(j->amplitude) && j->params.npeaks >= 1 &&
                          j->params.npeaks <= XPFC_MAX_PEAKS && j->params.dt > 0.0;
  if (!physics_ok) job.cls = -1;
  const int rc = v->sch->submit(job);
  if (rc) {
    for (auto& r : v->sch->take_results()) v->ready.push_back(r);
    return fail(rc, "job %lld rejected: %s", j->id,
                job.cls < 0 ? "no class for its grid or invalid physics"
                            : (j->steps < 1 ? "steps < 1" : "footprint exceeds the budget"));
  }
  v->store.specs[j->id] = *j;
  return XPFC_OK;
}

extern "C" int xpfc_service_poll(xpfc_service* v) {
  if (!v) return -fail(XPFC_ERR_ARG, "null service");
  const int r = v->sch->poll();
  for (auto& x : v->sch->take_results()) v->ready.push_back(x);
  return r;
}

extern "C" int xpfc_service_idle(const xpfc_service* v) { return v && v->sch->idle() ? 1 : 0; }

extern "C" int xpfc_service_run(xpfc_service* v, const xpfc_job* jobs, int n, int realtime) {
  if (!v || (n > 0 && !jobs) || n < 0) return fail(XPFC_ERR_ARG, "bad arguments");
  std::vector<const xpfc_job*> order;
  for (int i = 0; i < n; ++i) order.push_back(&jobs[i]);
  std::stable_sort(order.begin(), order.end(),
                   [](const xpfc_job* a, const xpfc_job* b) { return a->arrival < b->arrival; });
  size_t next = 0;
  while (next < order.size() || !v->sch->idle()) {
    const double now = v->clock();
    while (next < order.size() && (!realtime || order[next]->arrival <= now)) {
      xpfc_job j = *order[next++];
      if (!realtime || j.arrival < 0.0) j.arrival = now;
      const int rc = xpfc_service_submit(v, &j);
      if (rc && rc != XPFC_ERR_ARG) return rc;  // a rejected job is a result, not a failure
    }
    const int p = xpfc_service_poll(v);
    if (p < 0) return -p;
    if (p == 0) {
      // nothing to do until a block in flight ends or the next job arrives
      double wait = 5e-5;
      if (v->sch->idle() && next < order.size() && realtime)
        wait = std::max(0.0, std::min(1e-3, order[next]->arrival - v->clock()));
      std::this_thread::sleep_for(std::chrono::duration<double>(wait));
    }
  }
  return XPFC_OK;
}

extern "C" int xpfc_service_results(xpfc_service* v, xpfc_job_result* out, int max) {
  if (!v || (max > 0 && !out)) return -fail(XPFC_ERR_ARG, "bad arguments");
  const int n = std::min<int>(max, (int)v->ready.size());
  for (int i = 0; i < n; ++i) {
    const sched::Result& r = v->ready[(size_t)i];
    xpfc_job_result& o = out[i];
    o.id = r.id;
    o.status = r.status;
    o.steps_done = r.steps_done;
    o.F_ref = r.F_ref;
    o.f_mean = r.f_mean;
    o.mean = r.mean;
    o.t_arrival = r.t_arrival;
    o.t_admitted = r.t_admitted;
    o.t_finished = r.t_finished;
  }
  v->ready.erase(v->ready.begin(), v->ready.begin() + n);
  return n;
}

extern "C" long long xpfc_service_field(xpfc_service* v, long long id, double* host, long long capacity) {
  if (!v) return -fail(XPFC_ERR_ARG, "null service");
  auto it = v->store.fields.find(id);
  if (it == v->store.fields.end()) return -fail(XPFC_ERR_STATE, "no kept field for job %lld", id);
  const long long n = (long long)it->second.size();
  if (host) {
    if (capacity < n) return -fail(XPFC_ERR_ARG, "capacity %lld < %lld cells", capacity, n);
    std::copy(it->second.begin(), it->second.end(), host);
  }
  return n;
}

extern "C" int xpfc_service_get_stats(const xpfc_service* v, xpfc_service_stats* out) {
  if (!v || !out) return fail(XPFC_ERR_ARG, "null argument");
  const sched::Stats& s = v->sch->stats();
  out->submitted = s.submitted;
  out->rejected = s.rejected;
  out->done = s.done;
  out->nonfinite = s.nonfinite;
  out->queued = s.queued;
  out->running = s.running;
  out->blocks = s.blocks;
  out->queue_max = s.queue_max;
  out->resident_bytes = s.resident_bytes;
  out->peak_resident_bytes = s.peak_resident_bytes;
  out->cell_updates = s.cell_updates;
  out->elapsed = v->clock();
  return XPFC_OK;
}
