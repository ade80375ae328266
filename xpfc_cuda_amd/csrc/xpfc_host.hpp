// Private header of the host runtime (context layout, error helpers).
#pragma once
#include "../../include/xpfc.h"
#include "xpfc_internal.hpp"

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <utility>
#include <vector>

namespace xpfc {

// Records a message for xpfc_last_error() and returns `code`.
int fail(int code, const char* fmt, ...);

// Reference initial-condition helpers shared with the context code.
std::vector<double> linspace_ref(double x1, double x2, int N);
void fix_mean(std::vector<double>& psi, double psi0);
// legacy VTK structured points: dims[3], spacing[3], values in file order (x fastest)
int read_vtk_file(const char* path, int* dims, double* spacing, std::vector<double>* vals);
double c2hat_host(double kmag, const xpfc_params& p);

}  // namespace xpfc

#define HIPCHK(expr)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return xpfc::fail(XPFC_ERR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));   \
  } while (0)

#define LCHK(expr)                                                                          \
  do {                                                                                      \
    int r_ = (expr);                                                                        \
    if (r_ < 0) {                                                                           \
      return xpfc::fail(XPFC_ERR_DEVICE, "%s failed (%d): %s", #expr, r_,                   \
                        r_ == -2 ? hipGetErrorString(xpfc::launch_error())                  \
                                 : "variant or size not available");                        \
    }                                                                                       \
  } while (0)

// Per-instance state of a context.
struct XInst {
  xpfc_params p{};       // physics of this instance (grid equal to the context's)
  xpfc::InstArgs ia{};   // host copy of the device table entry
  double mean = 0.0;     // FP64 mean density (the k = 0 mode)
  double c2zero = 0.0;   // c2(k = 0), for the Parseval k = 0 term
  long long step = 0;
  bool have_field = false;
  std::vector<xpfc_sample> samples;
  double last_rmse = 0.0;
  bool have_rmse = false;
  bool nonfinite = false;           // a sampled sum of this instance was NaN / Inf
  bool nonfinite_reported = false;  // ... and xpfc_sync has reported it
  double* c2tab = nullptr;  // device C2 table (square / cubic grids, non-square 2-D boxes)
  int c2cap = 0;            // its capacity in entries
  xpfc::C2Row* c2rows = nullptr;  // device row index of the non-square 2-D table
  int c2rcap = 0;           // its capacity in rows
};

// One sampled step awaiting read-back: its row in the sums buffer and the
// instances that sample at it (with their step numbers).
struct XPending {
  int row;
  std::vector<std::pair<int, long long>> who;
};

struct xpfc_ctx {
  xpfc_params p{};  // instance 0's parameters: the grid of every instance
  int prec = 0;
  int n0 = 0, n1 = 0, n2 = 0, H = 0;
  long long N = 0, nmain = 0, nny = 0;
  long long xs = 0;     // plane stride of the main block (complex elements, >= n1*H)
  long long nyoff = 0;  // offset of the Nyquist plane = n0 * xs
  size_t csize = 0, rsize = 0;
  int K = 1;                 // instances
  long long stride = 0;      // complex elements per instance spectrum (nmain + nny)
  void* psi = nullptr;       // [K] packed spectra of psi - mean (state)
  void* work = nullptr;      // [K] nonlinear term / inverse-transform work fields
  void* aux = nullptr;       // [K] diagnostics / field I/O scratch
  void* tw0 = nullptr;       // exp(-2 pi i m / n0)
  void* tw1 = nullptr;       // exp(-2 pi i m / n1)
  void* twz = nullptr;       // exp(-2 pi i m / H)
  void* twp = nullptr;       // exp(-2 pi i k / n2), k < H
  double* px = nullptr;      // [K][npx_cap][XPART]
  double* pz = nullptr;      // [K][npz_cap][ZPART]
  double* sums = nullptr;    // [SUM_ROWS][K][NSUMS]
  xpfc::InstArgs* d_ia = nullptr;  // [K + 1]; entry K: scratch for operator calls
  int* d_ids = nullptr;      // [K + 1] = 0 .. K
  int* d_active = nullptr;   // [K] active instance list
  int* d_flags = nullptr;    // [K] divergence probe results
  int npx_cap = 0, npz_cap = 0;
  hipStream_t stream = nullptr;
  void* stage = nullptr;     // pinned host staging for field I/O (lazy)
  xpfc::KGeom g{};
  xpfc::Segs segX{}, segY{};
  int tc_x = 0, tc_y = 0, ntz = 256;  // pass variants
  bool nt_y = false;  // axis-1 passes with streaming hints (fields larger than the last-level cache)
  bool any_powf = false;  // some instance forms the FP32 powers (powf_fp32)
  bool tuned = false;
  bool use_graphs = true;
  std::vector<std::pair<std::pair<int, int>, hipGraphExec_t>> graphs;  // (steps, active) -> graph
  std::vector<XInst> inst;
  std::vector<int> active;
  std::vector<XPending> pending;
  std::vector<double> hsums;         // host rows for drain: [SUM_ROWS][K][NSUMS]
  bool nonfinite_unreported = false;  // an instance diverged since the last report
  int* d_budget = nullptr;           // [K] steps of the current block per slot (mixed contexts)
  std::vector<int> hbud;             // host copy of d_budget

  // ---- mixed-grid contexts (xpfc_create_mixed): instances of several grids.
  // Each grid is a class context holding that grid's instances (sharing this
  // context's stream); instance s is slot where[s].second of class where[s].first.
  std::vector<xpfc_ctx*> cls;
  std::vector<std::pair<int, int>> where;
  std::vector<std::vector<int>> inst_of;  // class k, slot i -> instance number
  bool borrowed_stream = false;      // a class context: the stream belongs to its owner
  struct RagLaunch {                 // one ragged launch: classes of one line family
    int grp = 0;                     // stream group: 0 on `stream`, 1 on `stream2`
    int fam = 1;                     // 1, 3 or 5 (2^k, 3*2^k, 5*2^k points)
    bool lut = false;                // axis 0: every class has a square / cubic box
    int first = 0, ncls = 0;         // its classes: d[first .. first + ncls)
    int nwg = 0;                     // workgroups
    size_t lds = 0;                  // dynamic LDS bytes
  };
  struct RagTables {                 // device class tables of one active-count configuration
    xpfc::RClass* d = nullptr;       // the classes of every launch, back to back
    std::vector<RagLaunch> x, y, z;  // axis 0, axis 1 (3-D classes), real space
    std::vector<int> nbx, nbz;       // per class: X / Z partial sums per instance
    bool two = false;                // launches on both streams
  };
  // Several-grid contexts step their instances in two groups, each class's
  // active list split in halves, on `stream` and `stream2`: the groups'
  // launches overlap, filling each other's ramp-up and tail.  Forked from and
  // joined back into `stream` around every block of steps.
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int mixed_groups = 2;              // 1 or 2 (XPFC_MIXED_STREAMS)
  std::map<std::vector<int>, RagTables> rtab;                  // key: active counts (+ budget flag)
  std::map<std::vector<int>, hipGraphExec_t> rgraphs;          // key: length, counts, flags
};
