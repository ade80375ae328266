// Job scheduler of the XPFC service: admission by device footprint, slots per
// grid class, blocks of steps with per-job budgets, the divergence stop and
// completion.  Plain C++ (no HIP): the device side is behind `Engine`, so the
// scheduling logic can be exercised against a stand-in engine on the host.
//
// The reference runs one simulation per process (main.cu:20-435) and has no
// scheduler; nothing here has a counterpart there.
#pragma once
#include <cstddef>
#include <deque>
#include <functional>
#include <vector>

namespace xpfc {
namespace sched {

enum JobStatus { kDone = 0, kNonfinite = 1, kRejected = 2 };

// A job as the scheduler sees it.
struct Job {
  long long id = 0;
  int cls = -1;            // grid class (index into the service's class list)
  long long steps = 0;     // steps to run
  double arrival = 0.0;    // seconds on the scheduler clock
  size_t footprint = 0;    // device bytes while resident
  long long cells = 0;     // cells updated per step
};

struct Result {
  long long id = 0;
  int status = kDone;
  long long steps_done = 0;
  double t_arrival = 0.0, t_admitted = 0.0, t_finished = 0.0;
  double F_ref = 0.0, f_mean = 0.0, mean = 0.0;  // filled by the engine
};

// The device side of one lane (a set of grid classes that advance together in
// the same blocks).  Class k has `slots[k]` slots; the scheduler passes
// per-class vectors over every class, filled only for the lane's classes.
// Every call is made between the lane's blocks except query().
class Engine {
 public:
  virtual ~Engine() = default;
  // The job's physics and initial field into slot `slot` of class k.
  virtual int start(const Job& job, int k, int slot) = 0;
  // Slots that step in the next blocks, per class.
  virtual int set_running(const std::vector<std::vector<int>>& slots) = 0;
  // One block of `len` steps; budget[k][slot] steps for each running slot.
  virtual int launch(int len, const std::vector<std::vector<int>>& budget) = 0;
  // 1: the block has finished, 0: still running, < 0: minus a status code.
  virtual int query() = 0;
  // finite[k][slot] = 0 once that slot's field has diverged.
  virtual int probe(std::vector<std::vector<int>>& finite) = 0;
  // Final diagnostics of a finished job (and its field, if kept); sets
  // r->status to kNonfinite when they are not finite.
  virtual int finish(const Job& job, int k, int slot, bool diverged, Result* r) = 0;
};

struct Config {
  std::vector<int> slots;     // per class
  std::vector<int> lane;      // per class: its lane (engine); empty = one lane per class
  double budget_bytes = 0.0;  // device bytes the resident jobs may hold
  int block = 50;             // steps per block while jobs wait
  int lookahead = 1;          // blocks per launch while no job waits
};

struct Stats {
  long long submitted = 0, rejected = 0, done = 0, nonfinite = 0;
  long long queued = 0, running = 0, blocks = 0, queue_max = 0;
  double resident_bytes = 0.0, peak_resident_bytes = 0.0;
  double cell_updates = 0.0;
};

class Scheduler {
 public:
  // engines[l] runs lane l (config.lane maps classes to lanes)
  Scheduler(Config cfg, std::vector<Engine*> engines, std::function<double()> clock);
  // Queue a job (XPFC_OK), or reject it: unknown class, fewer than 1 step, or
  // a footprint larger than the whole budget (XPFC_ERR_ARG, with a kRejected
  // result).
  int submit(const Job& job);
  // One round: for every lane whose block has finished, the divergence stop
  // and completions; then admission of queued jobs in arrival order into free
  // slots of lanes between blocks while memory allows (a job that does not
  // fit blocks the later ones; a job whose lane is mid-block waits without
  // blocking others); then the next block of every idle lane with running
  // jobs.  Returns 1 if anything happened, 0 if not, < 0 on failure.
  int poll();
  bool idle() const;
  std::vector<Result> take_results();
  const Stats& stats() const { return st_; }

 private:
  struct Run {
    Job job;
    long long left = 0;     // steps still to run
    long long taken = 0;    // steps run so far
    double admitted = 0.0;
    bool diverged = false;
    bool used = false;
  };
  struct Lane {
    Engine* eng = nullptr;
    std::vector<int> classes;
    bool in_flight = false;
    bool changed = true;   // its running set changed since the last block
  };
  int running_count() const;
  int lane_running(const Lane& l) const;
  int finish_block(Lane& l);
  int admit();
  int launch_block(Lane& l);

  Config cfg_;
  std::vector<Lane> lanes_;
  std::vector<int> lane_of_;                   // per class
  std::function<double()> clock_;
  std::deque<Job> queue_;                      // arrival order
  std::vector<std::vector<Run>> run_;          // [class][slot]
  std::vector<std::vector<int>> free_;         // free slots per class
  std::vector<std::vector<int>> budget_;       // steps of the block in flight
  std::vector<Result> results_;
  Stats st_;
};

}  // namespace sched
}  // namespace xpfc
