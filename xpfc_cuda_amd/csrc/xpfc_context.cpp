// The XPFC context: device fields of K simulations of one grid, twiddle and
// instance tables, the step driver (hipGraph step blocks), the autotuner,
// diagnostics, field I/O and the operator-level entry points.
//
// A context holds K >= 1 instances (xpfc_create: K = 1).  Instance s keeps its
// spectra at s * stride complex elements from the buffer bases; its physics
// (Phys, ZArgs) sits in a device table read by the kernels, so one launch per
// pass advances every active instance (blockIdx.y -> instance slot).
#include "xpfc_host.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <new>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

// Mixed-grid contexts: an instance call goes to the class context that holds
// the instance, with its slot there (named `cc` and `ss` in `expr`).
#define XPFC_ROUTE(c, s, expr)                                                            \
  do {                                                                                    \
    if ((c) && !(c)->cls.empty()) {                                                       \
      if ((s) < 0 || (s) >= (c)->K)                                                       \
        return fail(XPFC_ERR_ARG, "instance %d out of range [0, %d)", (int)(s), (c)->K);  \
      xpfc_ctx* const cc = (c)->cls[(size_t)(c)->where[(size_t)(s)].first];              \
      const int ss = (c)->where[(size_t)(s)].second;                                      \
      return expr;                                                                        \
    }                                                                                     \
  } while (0)
// Context-level operations a mixed-grid context does not offer.
#define XPFC_NOT_MIXED(c, what)                                                             \
  do {                                                                                      \
    if ((c) && !(c)->cls.empty())                                                           \
      return fail(XPFC_ERR_ARG, "%s is not available for a context of several grids", what); \
  } while (0)

using namespace xpfc;

namespace {

constexpr int SUM_ROWS = 1024;  // sampled steps held on the device per read-back
constexpr int DIAG_ROW = SUM_ROWS;  // extra row for on-demand diagnostics
constexpr int GRAPH_BLOCK = 50;  // steps per captured step block

// Grid extents the step passes compile: 8 .. 2048 points of the form f * 2^k
// with f in {1, 3, 5} and k >= 2 (so the half length n/2 of the contiguous
// axis is one of the compiled line lengths too).
bool supported_side(int n) {
  if (n < 8 || n > 2048) return false;
  int f = n;
  while (f % 2 == 0) f /= 2;
  return (f == 1 || f == 3 || f == 5) && n % (4 * f) == 0;
}

// Plane stride (complex elements) of the packed main block: planes of 64 KiB
// or more get 128 bytes of padding, so the axis-0 pass does not stride by a
// large power of two (measured: the 512^3 FP32 axis-0 pass 10 % faster).
// Fields of this many bytes per spectrum buffer (all instances) and more do
// not fit the 256 MiB last-level cache: their axis-1 passes use streaming
// hints.
constexpr size_t STREAM_BYTES = (size_t)256 << 20;

// Workgroup order over the XCDs (log2 of the runs of tiles dealt to each XCD
// in turn).  Measured against one contiguous run per XCD: the axis-0 pass with
// runs of 64 tiles -4 % at 512^3 FP32 and FP64, -6 % at 256^3 FP64, and with
// runs of 8 tiles -10 % on batches of 2-D 1024^2 fields (129 tiles each; no
// change at 512^2 and 256^2); the axis-1 passes with runs of two -6 % at 512^3
// FP32 but +3 % at 512^3 FP64 and +5 % at 256^3 FP64, so that order is a
// tuned variant (YRUNS_PAIRED).
constexpr int XCD_RUN_X = 6, XCD_RUN_Y = 1;

long long plane_stride(int n1, int H, size_t csize) {
  const long long plane = (long long)n1 * H;
  return plane * (long long)csize >= 65536 ? plane + (long long)(128 / csize) : plane;
}

// Runs f(i0, i1) over [0, n) on up to 16 host threads (field conversions at
// the host boundary).
template <class F>
void host_parallel(long long n, F f, long long grain = 1LL << 18) {
  const int nt = (int)std::min<long long>({16LL, std::max(1LL, n / grain),
                                           (long long)std::max(1u, std::thread::hardware_concurrency())});
  if (nt <= 1) {
    f(0LL, n);
    return;
  }
  std::vector<std::thread> th;
  const long long per = (n + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    const long long i0 = t * per, i1 = std::min(n, i0 + per);
    if (i0 < i1) th.emplace_back(f, i0, i1);
  }
  for (auto& x : th) x.join();
}

// FP64 sum in a fixed order: 256 contiguous blocks summed sequentially, then
// the block sums in order (independent of the thread count).
double fixed_order_sum(const double* v, long long n) {
  constexpr int NB = 256;
  double part[NB];
  // one thread per >= 2^18 elements, as for the conversions
  const long long grain = std::max(1LL, (long long)NB * (1LL << 18) / std::max(1LL, n));
  host_parallel(NB, [&](long long b0, long long b1) {
    for (long long b = b0; b < b1; ++b) {
      const long long i0 = n * b / NB, i1 = n * (b + 1) / NB;
      double s = 0.0;
      for (long long i = i0; i < i1; ++i) s += v[i];
      part[b] = s;
    }
  }, grain);
  double s = 0.0;
  for (int b = 0; b < NB; ++b) s += part[b];
  return s;
}

// Pinned host staging buffer of one real field in the working precision
// (allocated on first use, kept until the context is destroyed).
void* staging(xpfc_ctx* c) {
  if (!c->stage) {
    if (hipHostMalloc(&c->stage, (size_t)c->N * c->rsize, hipHostMallocDefault) != hipSuccess) {
      c->stage = nullptr;
      return nullptr;
    }
  }
  return c->stage;
}

template <typename T>
std::vector<T> twiddle_table(int L, int denom) {
  // tw[m] = exp(-2 pi i m / denom), m < L, evaluated in long double
  std::vector<T> t(2 * (size_t)L);
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for (int m = 0; m < L; ++m) {
    const long double ang = two_pi * (long double)m / (long double)denom;
    t[2 * m] = (T)std::cos(ang);
    t[2 * m + 1] = (T)(-std::sin(ang));
  }
  return t;
}

int upload_twiddles(xpfc_ctx* c, void** dst, int L, int denom) {
  const size_t bytes = (size_t)L * c->csize;
  HIPCHK(hipMalloc(dst, bytes));
  if (c->prec == 1) {
    auto t = twiddle_table<float>(L, denom);
    HIPCHK(hipMemcpy(*dst, t.data(), bytes, hipMemcpyHostToDevice));
  } else {
    auto t = twiddle_table<double>(L, denom);
    HIPCHK(hipMemcpy(*dst, t.data(), bytes, hipMemcpyHostToDevice));
  }
  return XPFC_OK;
}

// widest valid tile (16 columns FP32 / 8 FP64 keeps every row access 128 bytes;
// FP64 lines of 2048 points fit only 2 columns in LDS)
int default_tc(int prec, int L, const Segs& s) {
  const int pref = prec == 1 ? 16 : 8;
  for (int tc = pref; tc >= 2; tc /= 2)
    if (strided_blocks(prec, L, tc, s) >= 0) return tc;
  return -1;
}

// Captured step blocks bake in buffer pointers and pass variants (the active
// count is part of the cache key): drop them whenever one of those changes.
void drop_graphs(xpfc_ctx* c) {
  for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.second);
  c->graphs.clear();
}

void free_all(xpfc_ctx* c) {
  drop_graphs(c);
  for (auto& g : c->rgraphs) (void)hipGraphExecDestroy(g.second);
  c->rgraphs.clear();
  for (auto& t : c->rtab)
    if (t.second.d) (void)hipFree(t.second.d);
  c->rtab.clear();
  for (xpfc_ctx* k : c->cls) {
    free_all(k);
    delete k;
  }
  c->cls.clear();
  void* ptrs[] = {c->psi, c->work, c->aux, c->tw0, c->tw1, c->twz, c->twp, c->px,
                  c->pz, c->sums, c->d_ia, c->d_ids, c->d_active, c->d_flags, c->d_budget};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  for (XInst& in : c->inst) {
    if (in.c2tab) (void)hipFree(in.c2tab);
    if (in.c2rows) (void)hipFree(in.c2rows);
  }
  if (c->stream && !c->borrowed_stream) (void)hipStreamDestroy(c->stream);
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->stage) (void)hipHostFree(c->stage);
}

bool same_grid(const xpfc_params& a, const xpfc_params& b) {
  if (a.dim != b.dim || a.precision != b.precision) return false;
  for (int d = 0; d < a.dim; ++d)
    if (a.n[d] != b.n[d] || a.L[d] != b.L[d]) return false;
  return true;
}

// A peak term below exp(-ecut) is dropped: 1e-10 (FP32 fields) / 4e-18
// (FP64) of the denominator, which is >= 1.
double peak_ecut(int prec) { return prec == 1 ? 23.0 : 40.0; }

// |k|^2 beyond which every peak term of p is below exp(-ecut)
double peak_kcut2(const xpfc_params& p, double ecut) {
  double kc = 0.0;
  for (int i = 0; i < p.npeaks && i < XPFC_MAX_PEAKS; ++i)
    kc = std::max(kc, p.peak_k[i] + std::sqrt(ecut * 2.0 * p.peak_alpha[i] * p.peak_alpha[i]));
  return kc * kc;
}

// The largest m = |k|^2 / dk^2 of a square / cubic grid, axis extents (n0, n1, n2)
long long grid_mmax(int n0, int n1, int n2) {
  return (long long)(n0 / 2) * (n0 / 2) + (long long)(n1 / 2) * (n1 / 2) + (long long)(n2 / 2) * (n2 / 2);
}

// Entries of the C2 table over m of a square / cubic grid (up to the cutoff,
// or the largest m on the grid).
int c2_table_entries(double kcut2, double dksq, int n0, int n1, int n2) {
  return (int)std::min(std::ceil(kcut2 / dksq) + 1.0, (double)grid_mmax(n0, n1, n2) + 1.0);
}

// Rows of the C2 table of a 2-D grid whose box is not square: for each kz
// with k_z^2 below the cutoff, the modes |i| = 0 .. len-1, len one index past
// where the cutoff falls on the host, so that the device's own comparison
// (in the table builder, with the axis-0 pass's arithmetic) decides the edge
// and every mode outside the rows is beyond the cutoff.
std::vector<C2Row> c2_rows_2d(int n0, int H, double dk0, double dk2, double kcut2, long long* total) {
  std::vector<C2Row> rows;
  long long off = 0;
  const double kc = kcut2 * (1.0 + 1e-9);
  for (int l = 0; l <= H; ++l) {
    const double kz = dk2 * (double)l;
    const double rem = kc - kz * kz;
    if (rem <= 0.0) break;
    const long long len = std::min<long long>(n0 / 2 + 1, (long long)std::floor(std::sqrt(rem) / dk0) + 2);
    rows.push_back(C2Row{(int)off, (int)len});
    off += len;
  }
  *total = off;
  return rows;
}

// The state form (KGeom::xreal) of a grid with n1 points along axis 1 (1 in
// 2-D) and spectra of `sbytes` bytes.  Held along axis 0 in real space, the
// axis-0 pass writes one spectrum less per step and transforms the state once
// more: measured on MI355X, that pays for 3-D spectra of 16 MiB and more
// (256^3 FP32 +6 %, 512^3 FP32 +8 %, 128^3 FP64 +6 %, 512^3 FP64 +10 %) and
// costs below (128^3 FP32 -2 %, 64^3 FP32 -1 %) and for every 2-D grid, whose
// step has only two passes (256^2 to 2048^2: -5 to -13 %).  XPFC_STATE_FORM=0
// (psi_k) / 1 (axis-0 form) overrides the choice.
int state_form_of(int n1, size_t sbytes) {
  const char* e = std::getenv("XPFC_STATE_FORM");
  if (e && (e[0] == '0' || e[0] == '1') && e[1] == 0) return e[0] - '0';
  return (n1 > 1 && sbytes >= ((size_t)16 << 20)) ? 1 : 0;
}


// physics of one instance -> its kernel table entry
int build_inst(xpfc_ctx* c, const xpfc_params& p, XInst* in) {
  if (p.npeaks < 1 || p.npeaks > XPFC_MAX_PEAKS) return fail(XPFC_ERR_ARG, "npeaks out of range");
  if (!(p.dt > 0.0)) return fail(XPFC_ERR_ARG, "dt must be positive");
  in->p = p;
  Phys& ph = in->ia.ph;
  ph = Phys{};
  ph.dtM = p.dt * p.Mo;
  ph.one_m_eps = 1.0 - p.eps;
  ph.inv_dt = 1.0 / p.dt;
  ph.Mres = p.int_Mo_residual ? (double)(int)p.Mo : p.Mo;
  ph.npk = p.npeaks;
  ph.dcf_max = p.dcf_mode == XPFC_DCF_MAX;
  ph.ecut = peak_ecut(c->prec);
  for (int i = 0; i < p.npeaks; ++i) {
    ph.pk[i] = p.peak_k[i];
    ph.inv2a2[i] = 1.0 / (2.0 * p.peak_alpha[i] * p.peak_alpha[i]);
    ph.amp[i] = std::exp(-(p.sigma * p.peak_k[i]) * (p.sigma * p.peak_k[i]) / (2.0 * p.peak_rho_beta[i]));
  }
  ph.kcut2 = peak_kcut2(p, ph.ecut);
  in->c2zero = c2hat_host(0.0, p);
  if (c->g.cubic) {
    const int mcut = c2_table_entries(ph.kcut2, c->g.dksq, c->n0, c->n1, c->n2);
    if (mcut > in->c2cap) {
      HIPCHK(hipStreamSynchronize(c->stream));  // queued steps may still read the old table
      if (in->c2tab) (void)hipFree(in->c2tab);
      in->c2tab = nullptr;
      in->c2cap = 0;
      if (hipMalloc((void**)&in->c2tab, (size_t)mcut * sizeof(double)) != hipSuccess)
        return fail(XPFC_ERR_NOMEM, "C2 table allocation failed");
      in->c2cap = mcut;
    }
    ph.mcut = mcut;
    ph.c2tab = in->c2tab;
    LCHK(launch_c2table(in->c2tab, mcut, c->g.dksq, ph, c->stream));
  } else if (c->n1 == 1) {
    // 2-D box that is not square: the rows over kz (its columns)
    long long total = 0;
    const std::vector<C2Row> rows = c2_rows_2d(c->n0, c->H, c->g.dk0, c->g.dk2, ph.kcut2, &total);
    if (!rows.empty()) {
      HIPCHK(hipStreamSynchronize(c->stream));  // queued steps may still read the old rows
      if (total > in->c2cap) {
        if (in->c2tab) (void)hipFree(in->c2tab);
        in->c2tab = nullptr;
        in->c2cap = 0;
        if (hipMalloc((void**)&in->c2tab, (size_t)total * sizeof(double)) != hipSuccess)
          return fail(XPFC_ERR_NOMEM, "C2 table allocation failed");
        in->c2cap = (int)total;
      }
      if ((int)rows.size() > in->c2rcap) {
        if (in->c2rows) (void)hipFree(in->c2rows);
        in->c2rows = nullptr;
        in->c2rcap = 0;
        if (hipMalloc((void**)&in->c2rows, rows.size() * sizeof(C2Row)) != hipSuccess)
          return fail(XPFC_ERR_NOMEM, "C2 row allocation failed");
        in->c2rcap = (int)rows.size();
      }
      HIPCHK(hipMemcpy(in->c2rows, rows.data(), rows.size() * sizeof(C2Row), hipMemcpyHostToDevice));
      ph.c2tab = in->c2tab;
      ph.c2rows = in->c2rows;
      ph.c2nl = (int)rows.size();
      LCHK(launch_c2rows(in->c2tab, in->c2rows, ph.c2nl, c->g.dk0, c->g.dk2, ph, c->stream));
    }
  }
  in->ia.za.a = p.a;
  in->ia.za.b = p.b;
  in->ia.za.powf32 = p.powf_fp32 ? 1 : 0;
  in->ia.za.scale = 1.0 / (double)c->N;
  in->ia.za.mean = in->mean;
  // the real-space kernels carry the FP32-power form only while an instance
  // asks for it; captured step blocks bake the kernel choice in
  bool any = false;
  for (const XInst& x : c->inst) any = any || x.p.powf_fp32 != 0;
  if (any != c->any_powf) {
    c->any_powf = any;
    drop_graphs(c);
  }
  return XPFC_OK;
}

int upload_inst(xpfc_ctx* c, int s) {
  c->inst[s].ia.za.mean = c->inst[s].mean;
  HIPCHK(hipMemcpyAsync(c->d_ia + s, &c->inst[s].ia, sizeof(InstArgs), hipMemcpyHostToDevice,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return XPFC_OK;
}

Batch bt_of(const xpfc_ctx* c, const int* slots) {
  return Batch{slots, c->d_ia, c->stride, (long long)c->npx_cap * XPART,
               (long long)c->npz_cap * ZPART, nullptr, 0};
}
Batch bt_active(const xpfc_ctx* c) { return bt_of(c, c->d_active); }
Batch bt_one(const xpfc_ctx* c, int s) { return bt_of(c, c->d_ids + s); }

void* ny_of(const xpfc_ctx* c, void* spec) { return (char*)spec + (size_t)c->nyoff * c->csize; }
// real-space pass lines in the context's own layout (input and output)
ZLines zl_own(const xpfc_ctx* c) { return ZLines{c->n1, c->xs, c->xs}; }
void* inst_ptr(const xpfc_ctx* c, void* base, int s) {
  return (char*)base + (size_t)s * c->stride * c->csize;
}

// Reads the sums of pending sampled steps back into the sample logs (defined
// with the step driver); drain_report also reports newly diverged instances.
int drain(xpfc_ctx* c);
int drain_report(xpfc_ctx* c);
// Builds instance s's spectra from the fluctuation held in its aux region
// (defined with the field I/O).
int field_from_aux(xpfc_ctx* c, int s);

// Contexts of several grids (defined at the end of the file).
int mixed_step(xpfc_ctx* c, int nsteps);
int mixed_set_active(xpfc_ctx* c, const int* slots, int n);
int mixed_sync(xpfc_ctx* c);
int mixed_probe(xpfc_ctx* c, int* finite);

int check_inst(const xpfc_ctx* c, int s) {
  if (!c) return fail(XPFC_ERR_ARG, "null context");
  if (s < 0 || s >= c->K) return fail(XPFC_ERR_ARG, "instance %d out of range [0, %d)", s, c->K);
  return XPFC_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// creation
// ---------------------------------------------------------------------------
namespace {
// A context of K instances of one grid on its own stream, or (shared != null)
// on the stream of the mixed-grid context it is a class of.
int create_impl(const xpfc_params* ps, int K, hipStream_t shared, xpfc_ctx** out) {
  if (!ps || !out || K < 1) return fail(XPFC_ERR_ARG, "bad arguments");
  *out = nullptr;
  xpfc_params p = ps[0];
  if (p.dim != 2 && p.dim != 3) return fail(XPFC_ERR_ARG, "dim must be 2 or 3");
  if (p.dim == 2) {
    p.n[2] = 1;
    p.L[2] = 1.0;
  }
  const int nd = p.dim;
  for (int d = 0; d < nd; ++d) {
    if (!supported_side(p.n[d]))
      return fail(XPFC_ERR_ARG, "grid extent n[%d]=%d must be 2^k, 3*2^k or 5*2^k points in [8, 2048]", d,
                  p.n[d]);
    if (!(p.L[d] > 0.0)) return fail(XPFC_ERR_ARG, "box length L[%d] must be positive", d);
  }
  if (p.precision != XPFC_FP64 && p.precision != XPFC_FP32_MIXED)
    return fail(XPFC_ERR_ARG, "bad precision");
  for (int s = 1; s < K; ++s) {
    xpfc_params q = ps[s];
    if (q.dim == 2) {
      q.n[2] = 1;
      q.L[2] = 1.0;
    }
    if (!same_grid(p, q)) return fail(XPFC_ERR_ARG, "instance %d: grid/precision differ from instance 0", s);
  }

  xpfc_ctx* c = new (std::nothrow) xpfc_ctx();
  if (!c) return fail(XPFC_ERR_NOMEM, "host allocation failed");
  c->p = p;
  c->K = K;
  c->prec = p.precision == XPFC_FP32_MIXED ? 1 : 0;
  // axis roles: 2-D (Mx, My) -> (n0, n1, n2) = (Mx, 1, My); 3-D (Mx, My, Mz)
  double L0, L1, L2;
  if (nd == 2) {
    c->n0 = p.n[0]; c->n1 = 1; c->n2 = p.n[1];
    L0 = p.L[0]; L1 = 1.0; L2 = p.L[1];
  } else {
    c->n0 = p.n[0]; c->n1 = p.n[1]; c->n2 = p.n[2];
    L0 = p.L[0]; L1 = p.L[1]; L2 = p.L[2];
  }
  c->H = c->n2 / 2;
  c->N = (long long)c->n0 * c->n1 * c->n2;
  c->nmain = (long long)c->n0 * c->n1 * c->H;
  c->nny = (long long)c->n0 * c->n1;
  c->csize = c->prec == 1 ? 8 : 16;
  c->xs = plane_stride(c->n1, c->H, c->csize);
  c->nyoff = (long long)c->n0 * c->xs;
  c->stride = c->nyoff + c->nny;
  c->rsize = c->prec == 1 ? 4 : 8;

  int rc = XPFC_OK;
  auto bail = [&](int r) {
    free_all(c);
    delete c;
    return r;
  };
  if (hipSetDevice(p.device) != hipSuccess) return bail(fail(XPFC_ERR_DEVICE, "hipSetDevice(%d) failed", p.device));
  // the device address of the fixed scratch array of the allocation-free
  // entry point, looked up once per device with the device current; a failure
  // only disables that entry point, which reports it
  (void)noalloc_scratch_setup(p.device);
  if (shared) {
    c->stream = shared;
    c->borrowed_stream = true;
  } else if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    return bail(fail(XPFC_ERR_DEVICE, "stream creation failed"));
  }
  const size_t sbytes = (size_t)K * c->stride * c->csize;
  // the runtime may grant allocations beyond the device's memory and fail
  // later; refuse what cannot fit up front
  size_t mfree = 0, mtotal = 0;
  if (hipMemGetInfo(&mfree, &mtotal) == hipSuccess && sbytes > mfree / 3)
    return bail(fail(XPFC_ERR_NOMEM, "%d instance(s) need 3 x %zu bytes of device memory, %zu free", K,
                     sbytes, mfree));
  if (hipMalloc(&c->psi, sbytes) != hipSuccess || hipMalloc(&c->work, sbytes) != hipSuccess ||
      hipMalloc(&c->aux, sbytes) != hipSuccess)
    return bail(fail(XPFC_ERR_NOMEM, "device allocation of %zu bytes x3 failed", sbytes));
  if ((rc = upload_twiddles(c, &c->tw0, c->n0, c->n0)) != XPFC_OK) return bail(rc);
  if ((rc = upload_twiddles(c, &c->tw1, c->n1, c->n1)) != XPFC_OK) return bail(rc);
  if ((rc = upload_twiddles(c, &c->twz, c->H, c->H)) != XPFC_OK) return bail(rc);
  if ((rc = upload_twiddles(c, &c->twp, c->H, c->n2)) != XPFC_OK) return bail(rc);

  // strided segments: main half spectrum and the Nyquist plane
  c->segX.s[0] = Seg{0, 0, (int)c->xs, 1, (int)(c->n1 * c->H), 1, 0, 0};
  c->segX.s[1] = Seg{c->nyoff, 0, c->n1, 1, c->n1, 1, 0, 1};
  c->segY.s[0] = Seg{0, c->xs, c->H, 1, c->H, c->n0, 0, 0};
  c->segY.s[1] = Seg{c->nyoff, 0, 1, c->n1, c->n0, 1, 0, 1};
  c->segX.lgc = -1;  // set per launch from the tile count (x_segs)
  c->segY.lgc = -1;  // set per launch from the axis-1 variant
  c->tc_x = default_tc(c->prec, c->n0, c->segX);
  c->tc_y = c->n1 > 1 ? default_tc(c->prec, c->n1, c->segY) : 0;
  if (c->tc_x <= 0 || (c->n1 > 1 && c->tc_y <= 0)) return bail(fail(XPFC_ERR_ARG, "unsupported grid"));
  c->ntz = 256;
  c->nt_y = sbytes >= STREAM_BYTES;
  // partial buffers sized for the largest variant (narrowest tile, smallest block)
  c->npx_cap = 4096;
  for (int tc : {2, 4, 8, 16}) c->npx_cap = std::max(c->npx_cap, strided_blocks(c->prec, c->n0, tc, c->segX));
  c->npz_cap = 1;
  for (int nz : {64, 128, 256, 512})
    c->npz_cap = std::max(c->npz_cap, zpass_blocks(c->prec, c->H, nz, (long long)c->n0 * c->n1));
  if (hipMalloc((void**)&c->px, (size_t)K * c->npx_cap * XPART * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&c->pz, (size_t)K * c->npz_cap * ZPART * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&c->sums, (size_t)(SUM_ROWS + 1) * K * NSUMS * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&c->d_ia, (size_t)(K + 1) * sizeof(InstArgs)) != hipSuccess ||
      hipMalloc((void**)&c->d_ids, (size_t)(K + 1) * sizeof(int)) != hipSuccess ||
      hipMalloc((void**)&c->d_active, (size_t)K * sizeof(int)) != hipSuccess ||
      hipMalloc((void**)&c->d_flags, (size_t)K * sizeof(int)) != hipSuccess ||
      hipMalloc((void**)&c->d_budget, (size_t)K * sizeof(int)) != hipSuccess)
    return bail(fail(XPFC_ERR_NOMEM, "table allocation failed"));
  try {
    c->hsums.assign((size_t)SUM_ROWS * K * NSUMS, 0.0);  // read-back rows of drain
    c->hbud.assign((size_t)K, 0);
  } catch (const std::bad_alloc&) {
    return bail(fail(XPFC_ERR_NOMEM, "host allocation failed"));
  }

  c->g.n0 = c->n0;
  c->g.n1 = c->n1;
  c->g.H = c->H;
  c->g.xs = c->xs;
  c->g.nyoff = c->nyoff;
  c->g.dk0 = 2.0 * M_PI / L0;
  c->g.dk1 = nd == 3 ? 2.0 * M_PI / L1 : 0.0;
  c->g.dk2 = 2.0 * M_PI / L2;
  c->g.cubic = (c->g.dk0 == c->g.dk2 && (nd == 2 || c->g.dk1 == c->g.dk0)) ? 1 : 0;
  c->g.dksq = c->g.dk0 * c->g.dk0;
  c->g.inv_n0 = 1.0 / (double)c->n0;
  c->g.xreal = state_form_of(c->n1, (size_t)c->stride * c->csize);

  c->inst.resize(K);
  std::vector<InstArgs> ia(K + 1);
  for (int s = 0; s < K; ++s) {
    c->inst[s].mean = ps[s].psi0;
    if ((rc = build_inst(c, ps[s], &c->inst[s])) != XPFC_OK) return bail(rc);
    ia[s] = c->inst[s].ia;
  }
  ia[K] = ia[0];
  std::vector<int> ids(K + 1);
  for (int s = 0; s <= K; ++s) ids[s] = s;
  if (hipMemcpy(c->d_ia, ia.data(), ia.size() * sizeof(InstArgs), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->d_ids, ids.data(), ids.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->d_active, ids.data(), K * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(XPFC_ERR_DEVICE, "table upload failed"));
  c->active.assign(ids.begin(), ids.begin() + K);
  const char* eg = std::getenv("XPFC_GRAPHS");
  c->use_graphs = !(eg && eg[0] == '0');
  *out = c;
  return XPFC_OK;
}
}  // namespace

extern "C" int xpfc_create_batch(const xpfc_params* ps, int K, xpfc_ctx** out) {
  return create_impl(ps, K, nullptr, out);
}

extern "C" int xpfc_create(const xpfc_params* p, xpfc_ctx** out) { return xpfc_create_batch(p, 1, out); }

extern "C" size_t xpfc_instance_bytes(const xpfc_params* p) {
  // what xpfc_create_batch allocates per instance: three spectra, the partial
  // sums of the largest pass variants, the sampled-sum rows
  if (!p || (p->dim != 2 && p->dim != 3) || (p->precision != XPFC_FP64 && p->precision != XPFC_FP32_MIXED))
    return 0;
  const int prec = p->precision == XPFC_FP32_MIXED ? 1 : 0;
  const int n0 = p->n[0], n1 = p->dim == 3 ? p->n[1] : 1, n2 = p->dim == 3 ? p->n[2] : p->n[1];
  if (!supported_side(n0) || !supported_side(n2) || (p->dim == 3 && !supported_side(n1))) return 0;
  const int H = n2 / 2;
  if (H < 1) return 0;
  const size_t csize = prec == 1 ? 8 : 16;
  const long long xs = plane_stride(n1, H, csize);
  const long long nyoff = (long long)n0 * xs, nny = (long long)n0 * n1;
  Segs sx{};
  sx.s[0] = Seg{0, 0, (int)xs, 1, n1 * H, 1, 0, 0};
  sx.s[1] = Seg{nyoff, 0, n1, 1, n1, 1, 0, 1};
  int npx = 4096, npz = 1;
  for (int tc : {2, 4, 8, 16}) npx = std::max(npx, strided_blocks(prec, n0, tc, sx));
  for (int nz : {64, 128, 256, 512}) npz = std::max(npz, zpass_blocks(prec, H, nz, (long long)n0 * n1));
  // the instance's C2 table on square / cubic grids (as build_inst sizes it)
  size_t c2bytes = 0;
  const double L0 = p->L[0], L1 = p->dim == 3 ? p->L[1] : L0, L2 = p->dim == 3 ? p->L[2] : p->L[1];
  const double dk0 = 2.0 * M_PI / L0, dk1 = 2.0 * M_PI / L1, dk2 = 2.0 * M_PI / L2;
  if (dk0 == dk2 && dk1 == dk0 && p->npeaks >= 1 && p->npeaks <= XPFC_MAX_PEAKS) {
    c2bytes = (size_t)c2_table_entries(peak_kcut2(*p, peak_ecut(prec)), dk0 * dk0, n0, n1, n2) *
              sizeof(double);
  } else if (n1 == 1 && p->npeaks >= 1 && p->npeaks <= XPFC_MAX_PEAKS) {  // the 2-D rows
    long long total = 0;
    const size_t nrows = c2_rows_2d(n0, H, dk0, dk2, peak_kcut2(*p, peak_ecut(prec)), &total).size();
    c2bytes = (size_t)total * sizeof(double) + nrows * sizeof(C2Row);
  }
  return 3 * (size_t)(nyoff + nny) * csize + ((size_t)npx * XPART + (size_t)npz * ZPART) * sizeof(double) +
         (size_t)(SUM_ROWS + 1) * NSUMS * sizeof(double) + 2 * sizeof(InstArgs) + 3 * sizeof(int) + c2bytes;
}

extern "C" void xpfc_destroy(xpfc_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_all(c);
  delete c;
}

extern "C" int xpfc_get_params(const xpfc_ctx* c, xpfc_params* out) {
  if (!c || !out) return fail(XPFC_ERR_ARG, "null argument");
  if (!c->cls.empty()) return xpfc_get_inst_params(c, 0, out);
  *out = c->inst[0].p;
  return XPFC_OK;
}

extern "C" int xpfc_num_instances(const xpfc_ctx* c) { return c ? c->K : 0; }
// (mixed-grid contexts: those of instance 0's grid)
extern "C" long long xpfc_num_cells(const xpfc_ctx* c) {
  if (c && !c->cls.empty()) return xpfc_num_cells(c->cls[c->where[0].first]);
  return c ? c->N : 0;
}
extern "C" size_t xpfc_real_bytes(const xpfc_ctx* c) {
  if (c && !c->cls.empty()) return xpfc_real_bytes(c->cls[c->where[0].first]);
  return c ? (size_t)c->N * c->rsize : 0;
}
extern "C" size_t xpfc_spectrum_bytes(const xpfc_ctx* c) {
  if (c && !c->cls.empty()) return xpfc_spectrum_bytes(c->cls[c->where[0].first]);
  return c ? (size_t)c->n0 * c->n1 * (c->H + 1) * c->csize : 0;
}
extern "C" int xpfc_get_inst_params(const xpfc_ctx* c, int s, xpfc_params* out) {
  if (!c || !out || s < 0 || s >= c->K) return fail(XPFC_ERR_ARG, "bad arguments");
  if (!c->cls.empty()) return xpfc_get_inst_params(c->cls[c->where[s].first], c->where[s].second, out);
  *out = c->inst[s].p;
  return XPFC_OK;
}

extern "C" int xpfc_inst_set_params(xpfc_ctx* c, int s, const xpfc_params* p) {
  XPFC_ROUTE(c, s, xpfc_inst_set_params(cc, ss, p));
  int rc = check_inst(c, s);
  if (rc) return rc;
  if (!p) return fail(XPFC_ERR_ARG, "null params");
  xpfc_params q = *p;
  if (q.dim == 2) {
    q.n[2] = 1;
    q.L[2] = 1.0;
  }
  if (!same_grid(c->p, q)) return fail(XPFC_ERR_ARG, "grid/precision differ from the context's");
  if ((rc = drain(c)) != XPFC_OK) return rc;  // pending records use the old physics
  if ((rc = build_inst(c, q, &c->inst[s])) != XPFC_OK) return rc;
  return upload_inst(c, s);
}

extern "C" int xpfc_inst_reset(xpfc_ctx* c, int s) {
  XPFC_ROUTE(c, s, xpfc_inst_reset(cc, ss));
  int rc = check_inst(c, s);
  if (rc) return rc;
  if ((rc = drain(c)) != XPFC_OK) return rc;
  XInst& in = c->inst[s];
  in.step = 0;
  in.samples.clear();
  in.have_field = false;
  in.have_rmse = false;
  in.nonfinite = in.nonfinite_reported = false;
  return XPFC_OK;
}

extern "C" int xpfc_set_active(xpfc_ctx* c, const int* slots, int n) {
  if (!c) return fail(XPFC_ERR_ARG, "null context");
  if (!c->cls.empty()) return mixed_set_active(c, slots, n);
  std::vector<int> a;
  if (!slots) {
    for (int s = 0; s < c->K; ++s) a.push_back(s);
  } else {
    if (n < 0 || n > c->K) return fail(XPFC_ERR_ARG, "bad active count");
    std::vector<char> seen(c->K, 0);
    for (int i = 0; i < n; ++i) {
      if (slots[i] < 0 || slots[i] >= c->K || seen[slots[i]])
        return fail(XPFC_ERR_ARG, "bad or repeated instance %d", slots[i]);
      seen[slots[i]] = 1;
      a.push_back(slots[i]);
    }
  }
  if (!a.empty()) {
    HIPCHK(hipMemcpyAsync(c->d_active, a.data(), a.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  c->active = a;
  return XPFC_OK;
}

// ---------------------------------------------------------------------------
// transforms used for field I/O and diagnostics (one instance)
// ---------------------------------------------------------------------------
namespace {

// Axis-0 segments for tile width tc: runs of up to 2^XCD_RUN_X tiles, at
// least two rounds of them over the 8 XCDs; with fewer than 64 tiles (runs of
// 4), one contiguous run per XCD.
Segs x_segs(const xpfc_ctx* c, int tc) {
  Segs sg = c->segX;
  const int n = strided_blocks(c->prec, c->n0, tc, c->segX);
  int lg = XCD_RUN_X;
  while (lg >= 0 && n < (16 << lg)) --lg;
  sg.lgc = lg >= 2 ? lg : -1;
  return sg;
}

// axis-1 pass with variant v: tile width and workgroup order.  The streaming
// hints apply to tiles whose rows are whole 128-byte lines (measured at 512^3
// FP32: with 32- and 64-byte rows they double the pass's HBM traffic).
// `in`: the pass's input when it is not `spec` itself.
int launch_y(xpfc_ctx* c, void* spec, int v, int dir, const Batch& bt, int ny, const void* in = nullptr) {
  Segs sg = c->segY;
  sg.lgc = (v & YRUNS_PAIRED) ? XCD_RUN_Y : -1;
  const int tile = v & ~YRUNS_PAIRED;
  const bool nt = c->nt_y && (size_t)tile * c->csize >= 128;
  return launch_strided(c->prec, in ? in : spec, spec, sg, c->n1, tile, dir, c->tw1, bt, ny, nt, c->stream);
}

int forward_xy(xpfc_ctx* c, void* spec, const Batch& bt, int ny, bool do_x) {
  if (c->n1 > 1) LCHK(launch_y(c, spec, c->tc_y, -1, bt, ny));
  if (do_x)
    LCHK(launch_strided(c->prec, spec, x_segs(c, c->tc_x), c->n0, c->tc_x, -1, c->tw0, bt, ny, false, c->stream));
  return XPFC_OK;
}

int inverse_xy(xpfc_ctx* c, void* spec, const Batch& bt, int ny, bool do_x = true) {
  if (do_x)
    LCHK(launch_strided(c->prec, spec, x_segs(c, c->tc_x), c->n0, c->tc_x, +1, c->tw0, bt, ny, false, c->stream));
  if (c->n1 > 1) LCHK(launch_y(c, spec, c->tc_y, +1, bt, ny));
  return XPFC_OK;
}

// instance s: the state -> real field (working precision) in its aux region,
// optional real-space sums into its Z partials
int psi_to_real(xpfc_ctx* c, int s, bool sums) {
  const size_t sbytes = (size_t)c->stride * c->csize;
  HIPCHK(hipMemcpyAsync(inst_ptr(c, c->aux, s), inst_ptr(c, c->psi, s), sbytes,
                        hipMemcpyDeviceToDevice, c->stream));
  const Batch bt = bt_one(c, s);
  int rc = inverse_xy(c, c->aux, bt, 1, !c->g.xreal);
  if (rc) return rc;
  LCHK(launch_zpass(c->prec, Z_C2R, c->aux, ny_of(c, c->aux), c->aux, nullptr,
                    (long long)c->n0 * c->n1, c->H, zl_own(c), 256, bt, 1, c->twz, c->twp, c->pz,
                    sums, c->any_powf, c->stream));
  return XPFC_OK;
}

// fluctuation (host FP64) -> instance s's aux region (working precision)
int upload_fluct(xpfc_ctx* c, int s, const double* host, double mean) {
  const long long n = c->N;
  void* dst = inst_ptr(c, c->aux, s);
  void* st = staging(c);
  if (!st) return fail(XPFC_ERR_NOMEM, "pinned staging allocation failed");
  HIPCHK(hipStreamSynchronize(c->stream));  // the staging buffer may still feed a copy
  if (c->prec == 1) {
    float* f = (float*)st;
    host_parallel(n, [&](long long i0, long long i1) {
      for (long long i = i0; i < i1; ++i) f[i] = (float)(host[i] - mean);
    });
  } else {
    double* f = (double*)st;
    host_parallel(n, [&](long long i0, long long i1) {
      for (long long i = i0; i < i1; ++i) f[i] = host[i] - mean;
    });
  }
  // contiguous host planes -> the padded planes of the aux region
  const size_t row = (size_t)c->n1 * c->n2 * c->rsize;
  HIPCHK(hipMemcpy2DAsync(dst, (size_t)c->xs * c->csize, st, row, row, (size_t)c->n0,
                          hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return XPFC_OK;
}

}  // namespace

extern "C" int xpfc_inst_set_field(xpfc_ctx* c, int s, const double* host) {
  XPFC_ROUTE(c, s, xpfc_inst_set_field(cc, ss, host));
  int rc = check_inst(c, s);
  if (rc) return rc;
  if (!host) return fail(XPFC_ERR_ARG, "null field");
  if ((rc = drain(c)) != XPFC_OK) return rc;  // pending records use the old mean
  // Avg_calc.cu:3-18 mean, FP64 (fixed summation order)
  const double mean = fixed_order_sum(host, c->N) / (double)c->N;
  if (!std::isfinite(mean)) return fail(XPFC_ERR_NONFINITE, "field contains NaN/Inf");
  c->inst[s].mean = mean;
  if ((rc = upload_inst(c, s))) return rc;
  if ((rc = upload_fluct(c, s, host, mean))) return rc;
  return field_from_aux(c, s);
}

namespace {
// instance s's aux region holds the fluctuation psi - mean as reals: build the
// state spectrum and the nonlinear-term spectrum from it
int field_from_aux(xpfc_ctx* c, int s) {
  int rc = XPFC_OK;
  const long long nl = (long long)c->n0 * c->n1;
  const Batch bt = bt_one(c, s);
  // instance s's constants, pointers already at instance s's regions
  const Batch here{c->d_ids + s, c->d_ia, 0, 0, 0};
  void* aux = inst_ptr(c, c->aux, s);
  void* psi = inst_ptr(c, c->psi, s);
  void* work = inst_ptr(c, c->work, s);
  // psi_k = FFT(psi - mean) with the k = 0 slot set to zero; in the axis-0
  // state form, transformed back along axis 0
  LCHK(launch_zpass(c->prec, Z_R2C_FLUCT, aux, nullptr, psi, ny_of(c, psi), nl, c->H, zl_own(c), 256,
                    here, 1, c->twz, c->twp, nullptr, false, c->any_powf, c->stream));
  if ((rc = forward_xy(c, c->psi, bt, 1, true))) return rc;
  HIPCHK(hipMemsetAsync(psi, 0, c->csize, c->stream));
  if (c->g.xreal)
    LCHK(launch_strided(c->prec, c->psi, x_segs(c, c->tc_x), c->n0, c->tc_x, +1, c->tw0, bt, 1, false,
                        c->stream));
  // work = FFT_yz(N(psi)): the nonlinear term, all axes but axis 0
  LCHK(launch_zpass(c->prec, Z_R2C_NL, aux, nullptr, work, ny_of(c, work), nl, c->H, zl_own(c), 256,
                    here, 1, c->twz, c->twp, nullptr, false, c->any_powf, c->stream));
  if ((rc = forward_xy(c, c->work, bt, 1, false))) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  c->inst[s].have_field = true;
  return XPFC_OK;
}
}  // namespace

extern "C" int xpfc_inst_get_field(xpfc_ctx* c, int s, double* host) {
  XPFC_ROUTE(c, s, xpfc_inst_get_field(cc, ss, host));
  int rc = check_inst(c, s);
  if (rc) return rc;
  if (!host) return fail(XPFC_ERR_ARG, "null field");
  if (!c->inst[s].have_field) return fail(XPFC_ERR_STATE, "no field set");
  if ((rc = psi_to_real(c, s, false))) return rc;
  const size_t n = (size_t)c->N;
  const void* src = inst_ptr(c, c->aux, s);
  const size_t row = (size_t)c->n1 * c->n2 * c->rsize;  // one plane of reals on the host
  if (c->prec == 1) {
    void* st = staging(c);
    if (!st) return fail(XPFC_ERR_NOMEM, "pinned staging allocation failed");
    HIPCHK(hipMemcpy2DAsync(st, row, src, (size_t)c->xs * c->csize, row, (size_t)c->n0,
                            hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const float* f = (const float*)st;
    host_parallel((long long)n, [&](long long i0, long long i1) {
      for (long long i = i0; i < i1; ++i) host[i] = (double)f[i];
    });
  } else {
    HIPCHK(hipMemcpy2DAsync(host, row, src, (size_t)c->xs * c->csize, row, (size_t)c->n0,
                            hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return XPFC_OK;
}

extern "C" int xpfc_set_field(xpfc_ctx* c, const double* host) { return xpfc_inst_set_field(c, 0, host); }
extern "C" int xpfc_get_field(xpfc_ctx* c, double* host) { return xpfc_inst_get_field(c, 0, host); }

// ---------------------------------------------------------------------------
// initial conditions (initialize.cu:95-202)
// ---------------------------------------------------------------------------
extern "C" int xpfc_inst_initialize(xpfc_ctx* c, int s) {
  XPFC_ROUTE(c, s, xpfc_inst_initialize(cc, ss));
  int rc = check_inst(c, s);
  if (rc) return rc;
  const xpfc_params& p = c->inst[s].p;
  std::vector<double> psi((size_t)c->N, p.psi0);
  if (p.iconfig == 10) {
    return xpfc_inst_initialize_vtk(c, s, "initial_density_field.vtk");
  } else if (p.dim == 2) {
    const int Mx = p.n[0], My = p.n[1];
    const auto x = linspace_ref(0.0, p.L[0], Mx);
    const auto y = linspace_ref(0.0, p.L[1], My);
    const double a_tri = 2.0 / std::sqrt(3.0), a_sq = 1.0;
    const double k_sq = 2.0 * M_PI / a_sq, k_tri = 2.0 * M_PI / a_tri;
    for (int i = 0; i < Mx; ++i)
      for (int j = 0; j < My; ++j) {
        const double X = x[i], Y = y[j];
        double v;
        if (p.iconfig == 1)  // initialize.cu:96-99
          v = p.A_10 * ((3.0 * std::sqrt(2.0) - 2.0) * std::cos(k_sq * Y) - 2.0 * std::cos(k_sq * X)) -
              p.A_11 * std::cos(k_sq * X) * std::cos(k_sq * Y);
        else if (p.iconfig == 0)  // initialize.cu:116-118
          v = p.A_tri * (std::cos(k_tri * X) / 2.0 +
                         std::cos(k_tri * X / 2.0) * std::cos(std::sqrt(3.0) * k_tri * Y / 2.0));
        else
          v = 0.0;
        psi[(size_t)i * My + j] = v + p.psi0;
      }
    const double r = p.iconfig == 1 ? p.roa * a_sq : (p.iconfig == 0 ? p.roa * a_tri : 0.0);
    if (p.iseed == 1) {  // initialize.cu:158-191
      for (int i = 0; i < Mx; ++i)
        for (int j = 0; j < My; ++j) {
          double& v = psi[(size_t)i * My + j];
          const double dx = x[i] - p.h, dy = y[j] - p.k;
          if (p.ishape == 0) {
            if (std::sqrt(dx * dx + dy * dy) > r) v = p.psi0;
          } else {
            const double gs = std::exp(-(dx * dx + dy * dy) / p.roa);
            v = (v - p.psi0) * gs + p.psi0 * (1.0 - gs);
          }
        }
    }
  } else if (p.iconfig != 2) {
    return fail(XPFC_ERR_ARG, "3-D lattice initial conditions are not defined; use liquid (2) or noise");
  }
  fix_mean(psi, p.psi0);
  return xpfc_inst_set_field(c, s, psi.data());
}

extern "C" int xpfc_inst_initialize_vtk(xpfc_ctx* c, int s, const char* path) {
  XPFC_ROUTE(c, s, xpfc_inst_initialize_vtk(cc, ss, path));
  int rc = check_inst(c, s);
  if (rc) return rc;
  if (!path) return fail(XPFC_ERR_ARG, "null path");
  const xpfc_params& p = c->inst[s].p;
  int dims[3];
  double sp[3];
  std::vector<double> v;
  if ((rc = read_vtk_file(path, dims, sp, &v))) return rc;
  const int nz = p.dim == 3 ? p.n[2] : 1;
  if (dims[0] != p.n[0] || dims[1] != p.n[1] || dims[2] != nz)
    return fail(XPFC_ERR_ARG, "%s: dimensions %d %d %d do not match the grid %d %d %d", path, dims[0],
                dims[1], dims[2], p.n[0], p.n[1], nz);
  std::vector<double> psi((size_t)c->N);
  if (p.restart_order == 1) {
    if (p.dim != 2) return fail(XPFC_ERR_ARG, "restart_order 1 (reference fill) is 2-D only");
    psi = v;  // read_input.cu:208-218: psi[i*My + j] filled in file order
  } else {
    // file order is x fastest (write_vtk.cu:25-27); cells are [i][j][l], l fastest
    const int nx = dims[0], ny = dims[1];
    for (int l = 0; l < nz; ++l)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
          psi[((size_t)i * ny + j) * nz + l] = v[((size_t)l * ny + j) * nx + i];
  }
  fix_mean(psi, p.psi0);  // initialize.cu:200-202
  return xpfc_inst_set_field(c, s, psi.data());
}

// psi = psi0 + amplitude * randn(seed, cell) with its mean fixed to psi0, i.e.
// psi - psi0 = amplitude * (z - mean z): generated on the device with the
// draws of xpfc_randn, the mean of z summed in a fixed order
namespace {
// instance s's field psi0 + amplitude * (z - mean_z) from the draws' mean
// (both noise entry points); allocates nothing
int noise_field(xpfc_ctx* c, int s, double amplitude, unsigned long long seed, double mean_z) {
  if (!std::isfinite(mean_z)) return fail(XPFC_ERR_NONFINITE, "noise draws not finite");
  c->inst[s].mean = c->inst[s].p.psi0;
  int rc = upload_inst(c, s);
  if (rc) return rc;
  LCHK(launch_noise_fill(c->prec, inst_ptr(c, c->aux, s), c->N, (long long)c->n1 * c->n2, 2 * c->xs, seed,
                         amplitude, mean_z, c->stream));
  return field_from_aux(c, s);
}
}  // namespace

extern "C" int xpfc_inst_initialize_noise(xpfc_ctx* c, int s, double amplitude, unsigned long long seed) {
  XPFC_ROUTE(c, s, xpfc_inst_initialize_noise(cc, ss, amplitude, seed));
  int rc = check_inst(c, s);
  if (rc) return rc;
  if (!std::isfinite(amplitude)) return fail(XPFC_ERR_ARG, "non-finite amplitude");
  if ((rc = drain(c)) != XPFC_OK) return rc;  // pending records use the old mean
  double* part = c->px + (size_t)s * c->npx_cap * XPART;  // >= NOISE_PARTIALS doubles
  int nb = 0;
  LCHK(launch_noise_sum(c->N, seed, part, &nb, c->stream));
  std::vector<double> h(nb);
  HIPCHK(hipMemcpyAsync(h.data(), part, nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  double sz = 0.0;
  for (double v : h) sz += v;
  return noise_field(c, s, amplitude, seed, sz / (double)c->N);
}

// The same field with no allocation: the draws' partial sums go to the fixed
// device array of this GPU (its address looked up at context creation) and
// come back into a stack array; pending records of instance s are not read
// back (that would grow its sample log), the call refuses instead.
extern "C" int xpfc_inst_initialize_noise_noalloc(xpfc_ctx* c, int s, double amplitude,
                                                  unsigned long long seed) {
  XPFC_ROUTE(c, s, xpfc_inst_initialize_noise_noalloc(cc, ss, amplitude, seed));
  int rc = check_inst(c, s);
  if (rc) return rc;
  if (!std::isfinite(amplitude)) return fail(XPFC_ERR_ARG, "non-finite amplitude");
  for (const XPending& pr : c->pending)
    for (const auto& w : pr.who)
      if (w.first == s)
        return fail(XPFC_ERR_STATE, "sampled records of instance %d are pending: call xpfc_sync first", s);
  const int nb = noise_partials(c->N);
  if (nb < 1 || nb > NOISE_PARTIALS)
    return fail(XPFC_ERR_ARG, "%d partial sums exceed the fixed scratch of %d", nb, NOISE_PARTIALS);
  double* scr = noalloc_scratch(c->p.device);
  std::mutex* mu = noalloc_scratch_lock(c->p.device);
  if (!scr || !mu) return fail(XPFC_ERR_DEVICE, "no scratch array address for device %d", c->p.device);
  double h[NOISE_PARTIALS];
  {
    std::lock_guard<std::mutex> lk(*mu);  // one user of this GPU's array at a time
    int nbl = 0;
    LCHK(launch_noise_sum(c->N, seed, scr, &nbl, c->stream));
    if (nbl != nb) return fail(XPFC_ERR_DEVICE, "noise partial count %d != %d", nbl, nb);
    HIPCHK(hipMemcpyAsync(h, scr, (size_t)nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  double sz = 0.0;
  for (int i = 0; i < nb; ++i) sz += h[i];
  return noise_field(c, s, amplitude, seed, sz / (double)c->N);
}

extern "C" int xpfc_initialize(xpfc_ctx* c) { return xpfc_inst_initialize(c, 0); }
extern "C" int xpfc_initialize_noise(xpfc_ctx* c, double amplitude, unsigned long long seed) {
  return xpfc_inst_initialize_noise(c, 0, amplitude, seed);
}

// ---------------------------------------------------------------------------
// the step driver
// ---------------------------------------------------------------------------
namespace {

bool inst_samples(const XInst& in, long long n) {
  const long long ts = in.p.t_sample;
  return ts > 0 && (n % ts == 0 || (n + 1) % ts == 0);
}

// The launches of one step (four in 3-D, two in 2-D) over every active
// instance.  With `ev`, an event is recorded after each pass for timing.
// With `budget`, this is step `sib` of a budgeted block: instances whose
// budget is used up skip it.
int launch_one_step(xpfc_ctx* c, bool sample, int row, hipEvent_t* ev = nullptr, const int* budget = nullptr,
                    int sib = 0) {
  const long long nl = (long long)c->n0 * c->n1;
  Batch bt = bt_active(c);
  bt.budget = budget;
  bt.sib = sib;
  const int ny = (int)c->active.size();
  const int nbx = launch_xstep(c->prec, c->work, c->psi, x_segs(c, c->tc_x), c->n0, c->tc_x, c->tw0, c->g, bt, ny,
                               c->px, sample, c->stream);
  LCHK(nbx);
  if (ev) HIPCHK(hipEventRecord(ev[0], c->stream));
  // in the axis-0 state form the state itself feeds the next pass (axis 1 in
  // 3-D, real space in 2-D), which writes the working buffer
  if (c->n1 > 1) LCHK(launch_y(c, c->work, c->tc_y, +1, bt, ny, c->g.xreal ? c->psi : nullptr));
  if (ev) HIPCHK(hipEventRecord(ev[1], c->stream));
  void* zin = (c->n1 > 1 || !c->g.xreal) ? c->work : c->psi;
  LCHK(launch_zpass(c->prec, Z_STEP, zin, ny_of(c, zin), c->work, ny_of(c, c->work), nl, c->H,
                    zl_own(c), c->ntz, bt, ny, c->twz, c->twp, c->pz, sample, c->any_powf, c->stream));
  if (ev) HIPCHK(hipEventRecord(ev[2], c->stream));
  if (c->n1 > 1) LCHK(launch_y(c, c->work, c->tc_y, -1, bt, ny));
  if (ev) HIPCHK(hipEventRecord(ev[3], c->stream));
  if (sample)
    LCHK(launch_finalize(c->px, nbx, c->pz, zpass_blocks(c->prec, c->H, c->ntz, nl), bt, ny,
                         c->sums + (size_t)row * c->K * NSUMS, c->stream));
  return XPFC_OK;
}

// Unsampled steps run as replays of captured hipGraphs of GRAPH_BLOCK steps
// or a power of two below it, cached per context by (length, number of active
// instances): the kernels read the active list from device memory, so a new
// list of the same length reuses the graphs.  The captured work is the same
// launch sequence as an eager step, so the fields are identical.  Budgeted
// blocks (per-instance budgets read from d_budget) are cached under -len.
int plain_graph(xpfc_ctx* c, int len, hipGraphExec_t* out, bool budgeted = false) {
  const int ny = (int)c->active.size();
  const int key = budgeted ? -len : len;
  for (auto& g : c->graphs)
    if (g.first.first == key && g.first.second == ny) {
      *out = g.second;
      return XPFC_OK;
    }
  if (c->graphs.size() >= 32) drop_graphs(c);
  hipGraph_t graph = nullptr;
  HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rc = XPFC_OK;
  for (int i = 0; i < len && rc == XPFC_OK; ++i)
    rc = launch_one_step(c, false, 0, nullptr, budgeted ? c->d_budget : nullptr, i);
  const hipError_t ec = hipStreamEndCapture(c->stream, &graph);
  if (rc != XPFC_OK || ec != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc != XPFC_OK ? rc : fail(XPFC_ERR_DEVICE, "step capture failed: %s", hipGetErrorString(ec));
  }
  hipGraphExec_t ex = nullptr;
  const hipError_t ei = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) return fail(XPFC_ERR_DEVICE, "graph instantiation failed: %s", hipGetErrorString(ei));
  c->graphs.push_back({{key, ny}, ex});
  *out = ex;
  return XPFC_OK;
}

void advance(xpfc_ctx* c, long long n) {
  for (int s : c->active) c->inst[s].step += n;
}

// One block of `len` steps in which active instance s takes hbud[s] of them
// (one-grid contexts; the kernels of steps past an instance's budget return
// at once): the budgets go to the device, then the block runs as one
// captured graph.
int uniform_budget_block(xpfc_ctx* c, int len) {
  HIPCHK(hipMemcpyAsync(c->d_budget, c->hbud.data(), c->hbud.size() * sizeof(int), hipMemcpyHostToDevice,
                        c->stream));
  if (!c->use_graphs) {
    for (int i = 0; i < len; ++i)
      if (int rc = launch_one_step(c, false, 0, nullptr, c->d_budget, i)) return rc;
  } else {
    hipGraphExec_t ex = nullptr;
    if (int rc = plain_graph(c, len, &ex, true)) return rc;
    HIPCHK(hipGraphLaunch(ex, c->stream));
  }
  for (int s : c->active) c->inst[s].step += c->hbud[(size_t)s];
  return XPFC_OK;
}

int run_plain_steps(xpfc_ctx* c, long long run) {
  while (run > 0) {
    if (!c->use_graphs) {
      int rc = launch_one_step(c, false, 0);
      if (rc) return rc;
      advance(c, 1);
      --run;
      continue;
    }
    int len = GRAPH_BLOCK;
    if (run < GRAPH_BLOCK) {
      len = 1;
      while (2 * len <= run) len *= 2;
    }
    hipGraphExec_t ex = nullptr;
    int rc = plain_graph(c, len, &ex);
    if (rc) return rc;
    HIPCHK(hipGraphLaunch(ex, c->stream));
    advance(c, len);
    run -= len;
  }
  return XPFC_OK;
}

double ff_factor(const xpfc_params& p) {  // initialize.cu:63
  double ff = 1.0;
  for (int d = 0; d < p.dim; ++d) ff *= p.L[d] / (double)(p.n[d] - 1);
  return ff;
}

// sum f = 1/2 s2 - a/3 s3 + b/4 s4 - 1/2 sum_x psi (C2*psi)  (FE_calc.cu:64), with
// sum_x psi (C2*psi) = (1/N) sum_k c2 |psi_k|^2 by Parseval, k = 0 carried in FP64
double fe_sum(const xpfc_ctx* c, const XInst& in, double conv, double s2, double s3, double s4) {
  const double N = (double)c->N;
  const double conv_x = (conv + in.c2zero * (N * in.mean) * (N * in.mean)) / N;
  return 0.5 * s2 - (in.p.a / 3.0) * s3 + (in.p.b / 4.0) * s4 - 0.5 * conv_x;
}

void add_fe_record(xpfc_ctx* c, XInst& in, long long step, const double* s) {
  const double sf = fe_sum(c, in, s[0], s[3], s[4], s[5]);
  const double N = (double)c->N;
  if (!in.samples.empty() && in.samples.back().step == step && !in.samples.back().has_fe) {
    xpfc_sample& b = in.samples.back();
    b.has_fe = 1;
    b.F_ref = ff_factor(in.p) * sf;
    b.f_mean = sf / N;
    b.mean = s[2] / N;
    return;
  }
  xpfc_sample r{};
  r.step = step;
  r.has_fe = 1;
  r.F_ref = ff_factor(in.p) * sf;
  r.f_mean = sf / N;
  r.mean = s[2] / N;
  r.rmse = NAN;
  in.samples.push_back(r);
}

void add_sse_record(xpfc_ctx* c, XInst& in, long long step, double sse) {
  const double rmse = std::sqrt(sse) / (double)c->N;  // sqrt(sum_x R^2 / N), Parseval
  in.last_rmse = rmse;
  in.have_rmse = true;
  if (!in.samples.empty() && in.samples.back().step == step) {
    in.samples.back().has_sse = 1;
    in.samples.back().rmse = rmse;
    return;
  }
  xpfc_sample r{};
  r.step = step;
  r.has_sse = 1;
  r.rmse = rmse;
  r.F_ref = r.f_mean = r.mean = NAN;
  in.samples.push_back(r);
}

// Non-finite sums are recorded as they are (NaN / Inf in the records) and
// flagged per instance; the run of the other instances goes on.  xpfc_sync
// reports each newly diverged instance once (XPFC_ERR_NONFINITE).
int drain(xpfc_ctx* c) {
  if (c->pending.empty()) return XPFC_OK;
  const int rows = (int)c->pending.size();
  double* h = c->hsums.data();  // sized for SUM_ROWS rows at creation
  HIPCHK(hipMemcpyAsync(h, c->sums, (size_t)rows * c->K * NSUMS * sizeof(double),
                        hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (const XPending& pr : c->pending)
    for (const auto& w : pr.who) {
      XInst& in = c->inst[w.first];
      const double* s = &h[((size_t)pr.row * c->K + w.first) * NSUMS];
      bool bad = false;
      for (int k = 0; k < NSUMS; ++k) bad |= !std::isfinite(s[k]);
      if (bad && !in.nonfinite) {
        in.nonfinite = true;
        c->nonfinite_unreported = true;
      }
      const long long n = w.second, ts = in.p.t_sample;
      if (n % ts == 0) add_sse_record(c, in, n, s[1]);
      if ((n + 1) % ts == 0) add_fe_record(c, in, n + 1, s);
    }
  c->pending.clear();
  return XPFC_OK;
}

// drain, then report instances that diverged since the last report
int drain_report(xpfc_ctx* c) {
  int rc = drain(c);
  if (rc) return rc;
  if (c->nonfinite_unreported) {
    c->nonfinite_unreported = false;
    std::string who;
    for (int s = 0; s < c->K; ++s)
      if (c->inst[s].nonfinite && !c->inst[s].nonfinite_reported) {
        c->inst[s].nonfinite_reported = true;
        if (who.size() < 200) who += (who.empty() ? "" : " ") + std::to_string(s);
      }
    return fail(XPFC_ERR_NONFINITE, "non-finite value in a sampled reduction (instance %s)", who.c_str());
  }
  return XPFC_OK;
}

int fe_now(xpfc_ctx* c, int s, double* F, double* fm, double* mean_rs);

// main.cu:286: the FE record of the initial field of every active instance
// that samples and has not stepped yet
int initial_records(xpfc_ctx* c) {
  for (int s : c->active) {
    XInst& in = c->inst[s];
    if (in.step == 0 && in.p.t_sample > 0 && in.samples.empty()) {
      double F, fm, mrs;
      int rc = fe_now(c, s, &F, &fm, &mrs);
      if (rc) return rc;
      xpfc_sample r{};
      r.step = 0;
      r.has_fe = 1;
      r.F_ref = F;
      r.f_mean = fm;
      r.mean = mrs;
      r.rmse = NAN;
      in.samples.push_back(r);
    }
  }
  return XPFC_OK;
}

int fe_now(xpfc_ctx* c, int s, double* F, double* fm, double* mean_rs) {
  int nb = 0;
  const Batch bt = bt_one(c, s);
  // psi_k: the state, or in the axis-0 form psi_k as the axis-0 pass forms
  // it (FFT_0(state) / n0, k = 0 slot zero) in the aux region
  void* spec = c->psi;
  if (c->g.xreal) {
    void* aux = inst_ptr(c, c->aux, s);
    HIPCHK(hipMemcpyAsync(aux, inst_ptr(c, c->psi, s), (size_t)c->stride * c->csize, hipMemcpyDeviceToDevice,
                          c->stream));
    LCHK(launch_strided(c->prec, c->aux, x_segs(c, c->tc_x), c->n0, c->tc_x, -1, c->tw0, bt, 1, false,
                        c->stream));
    LCHK(launch_scale(c->prec, aux, 2LL * c->stride, c->g.inv_n0, c->stream));
    HIPCHK(hipMemsetAsync(aux, 0, c->csize, c->stream));
    spec = c->aux;
  }
  LCHK(launch_kspace_sums(c->prec, spec, c->nmain, c->nny, c->g, bt, c->px, &nb, c->stream));
  int rc = psi_to_real(c, s, true);
  if (rc) return rc;
  double* row = c->sums + (size_t)DIAG_ROW * c->K * NSUMS;
  LCHK(launch_finalize(c->px, nb, c->pz, zpass_blocks(c->prec, c->H, 256, (long long)c->n0 * c->n1), bt,
                       1, row, c->stream));
  double h[NSUMS];
  HIPCHK(hipMemcpyAsync(h, row + (size_t)s * NSUMS, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (double v : h)
    if (!std::isfinite(v)) return fail(XPFC_ERR_NONFINITE, "non-finite free-energy sum");
  const XInst& in = c->inst[s];
  const double sf = fe_sum(c, in, h[0], h[3], h[4], h[5]);
  *F = ff_factor(in.p) * sf;
  *fm = sf / (double)c->N;
  *mean_rs = h[2] / (double)c->N;
  return XPFC_OK;
}

// ---------------------------------------------------------------------------
// autotuner: times every variant of each pass on scratch buffers and keeps the
// fastest; results are cached per (precision, grid, instances) for the
// process.  The variants only regroup lines into workgroups or change register
// allocation, so the choice never changes the computed field.
// ---------------------------------------------------------------------------
std::mutex g_tune_mu;
std::map<std::array<int, 6>, std::array<int, 3>> g_tune_cache;

int time_pass(xpfc_ctx* c, int which, int v, void* a, void* b, int ny, int reps, float* ms) {
  const long long nl = (long long)c->n0 * c->n1;
  const Batch bt = bt_of(c, c->d_ids);
  auto launch = [&]() -> int {
    if (which == 0)
      return launch_xstep(c->prec, a, b, x_segs(c, v), c->n0, v, c->tw0, c->g, bt, ny, c->px, false, c->stream);
    if (which == 1) return launch_y(c, a, v, -1, bt, ny);
    return launch_zpass(c->prec, Z_STEP, a, ny_of(c, a), a, ny_of(c, a), nl, c->H, zl_own(c), v, bt, ny,
                        c->twz, c->twp, c->pz, false, c->any_powf, c->stream);
  };
  int r = launch();  // warm-up; also tells whether the variant exists
  if (r == -3) return -3;
  if (r < 0) return fail(XPFC_ERR_DEVICE, "autotune launch failed (%d)", r);
  struct Ev {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Ev() {
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
    }
  } ev;
  HIPCHK(hipEventCreate(&ev.e0));
  HIPCHK(hipEventCreate(&ev.e1));
  // average over n launches replayed from a captured graph, as the step runs
  // them: launched one by one from the host, kernels of a few microseconds
  // (small grids) are all timed at the launch rate and cannot be told apart
  auto timed = [&](int n, float* out) -> int {
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int lr = 0;
    for (int i = 0; i < n && lr >= 0; ++i) lr = launch();
    const hipError_t ec = hipStreamEndCapture(c->stream, &graph);
    if (lr < 0 || ec != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return fail(XPFC_ERR_DEVICE, "autotune capture failed");
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t ei = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) return fail(XPFC_ERR_DEVICE, "autotune graph: %s", hipGetErrorString(ei));
    struct Exec {
      hipGraphExec_t e;
      ~Exec() { (void)hipGraphExecDestroy(e); }
    } hold{ex};
    HIPCHK(hipGraphLaunch(ex, c->stream));  // warm-up replay
    HIPCHK(hipEventRecord(ev.e0, c->stream));
    HIPCHK(hipGraphLaunch(ex, c->stream));
    HIPCHK(hipEventRecord(ev.e1, c->stream));
    HIPCHK(hipEventSynchronize(ev.e1));
    HIPCHK(hipEventElapsedTime(out, ev.e0, ev.e1));
    *out /= (float)n;
    return 0;
  };
  float t1 = 0.f;
  if ((r = timed(reps, &t1))) return r;
  // about 4 ms of launches per measurement, best of two
  const int n = std::max(reps, std::min(64, (int)(4.0f / std::max(t1, 1e-3f))));
  float ta = 0.f, tb = 0.f;
  if ((r = timed(n, &ta)) || (r = timed(n, &tb))) return r;
  *ms = std::min(ta, tb);
  return 0;
}

int autotune(xpfc_ctx* c) {
  // tune with up to 8 instances side by side (the batched launch shape)
  const int ny = std::min(c->K, 8);
  const std::array<int, 6> key{c->prec, c->n0, c->n1, c->n2, ny, c->nt_y ? 1 : 0};
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      c->tc_x = it->second[0];
      c->tc_y = it->second[1];
      c->ntz = it->second[2];
      c->tuned = true;
      drop_graphs(c);
      return XPFC_OK;
    }
  }
  const size_t sbytes = (size_t)ny * c->stride * c->csize;
  // two scratch spectra per tuned instance, when free memory holds them with
  // room to spare; otherwise (the largest grids, whose three buffers fill most
  // of the device) the variants are timed in place in the context's
  // diagnostics buffer, used as both operands: its contents are rebuilt by
  // every call that reads them, and timing does not depend on the values
  void *a = nullptr, *b = nullptr;
  size_t mfree = 0, mtotal = 0;
  const bool room = hipMemGetInfo(&mfree, &mtotal) == hipSuccess && 2 * sbytes + (mtotal >> 4) <= mfree;
  if (room && (hipMalloc(&a, sbytes) != hipSuccess || hipMalloc(&b, sbytes) != hipSuccess)) {
    if (a) (void)hipFree(a);
    a = b = nullptr;
    (void)hipGetLastError();  // a refused allocation is not a launch error
  }
  const bool own = a && b;
  if (!own) a = b = c->aux;
  int rc = XPFC_OK;
  if (hipMemsetAsync(a, 0, sbytes, c->stream) != hipSuccess ||
      hipMemsetAsync(b, 0, sbytes, c->stream) != hipSuccess)
    rc = fail(XPFC_ERR_DEVICE, "autotune scratch clear failed");
  const int reps = 2;  // first estimate; time_pass then sizes its measurement
  int best[3] = {c->tc_x, c->tc_y, c->ntz};
  // axis-0 and axis-1: tile width; real space: block (+ZFREE_REGS).  Listed in
  // order of preference (wide tiles write whole 128-byte lines); a later
  // variant must be 3 % faster to replace an earlier one.
  const std::vector<int> xs{16, 8, 4, 2};
  const std::vector<int> ys{16, 16 | YRUNS_PAIRED, 8, 8 | YRUNS_PAIRED, 4, 4 | YRUNS_PAIRED, 2, 2 | YRUNS_PAIRED};
  const std::vector<int> zs{256, 256 | ZFREE_REGS, 128, 128 | ZFREE_REGS, 512, 512 | ZFREE_REGS,
                            64, 64 | ZFREE_REGS};
  for (int which = 0; which < 3 && rc == XPFC_OK; ++which) {
    if (which == 1 && c->n1 <= 1) continue;
    const std::vector<int>& vs = which == 0 ? xs : which == 1 ? ys : zs;
    // three rounds over the variants, each variant's best time kept: clock and
    // cache drift during tuning hits every variant alike
    std::vector<float> tbest(vs.size(), 1e30f);
    for (int round = 0; round < 3 && rc == XPFC_OK; ++round) {
      for (size_t i = 0; i < vs.size(); ++i) {
        float ms = 0.f;
        const int r = time_pass(c, which, vs[i], a, b, ny, reps, &ms);
        if (r == -3) continue;
        if (r) {
          rc = r;
          break;
        }
        tbest[i] = std::min(tbest[i], ms);
      }
    }
    float bt = 1e30f;
    for (size_t i = 0; i < vs.size(); ++i)
      if (tbest[i] < 0.97f * bt) {
        bt = tbest[i];
        best[which] = vs[i];
      }
  }
  (void)hipStreamSynchronize(c->stream);
  if (own) {
    (void)hipFree(a);
    (void)hipFree(b);
  }
  if (rc) return rc;
  c->tc_x = best[0];
  c->tc_y = best[1];
  c->ntz = best[2];
  c->tuned = true;
  drop_graphs(c);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune_cache[key] = {best[0], best[1], best[2]};
  return XPFC_OK;
}

bool autotune_enabled() {
  const char* e = std::getenv("XPFC_AUTOTUNE");
  return !(e && e[0] == '0');
}

}  // namespace

extern "C" int xpfc_autotune(xpfc_ctx* c, int* variants) {
  if (!c) return fail(XPFC_ERR_ARG, "null context");
  if (!c->cls.empty()) {  // one launch shape per pass for every grid: nothing to choose
    if (variants) variants[0] = variants[1] = variants[2] = 0;
    return XPFC_OK;
  }
  int rc = autotune(c);
  if (rc) return rc;
  if (variants) {
    variants[0] = c->tc_x;
    variants[1] = c->tc_y;
    variants[2] = c->ntz;
  }
  return XPFC_OK;
}

extern "C" int xpfc_set_variants(xpfc_ctx* c, const int* v) {
  if (!c || !v) return fail(XPFC_ERR_ARG, "null argument");
  XPFC_NOT_MIXED(c, "xpfc_set_variants");
  const long long nl = (long long)c->n0 * c->n1;
  if (v[0] > 0 && strided_blocks(c->prec, c->n0, v[0], c->segX) < 0)
    return fail(XPFC_ERR_ARG, "tile width %d not available for n0 = %d", v[0], c->n0);
  if (c->n1 > 1 && v[1] > 0 && strided_blocks(c->prec, c->n1, v[1] & ~YRUNS_PAIRED, c->segY) < 0)
    return fail(XPFC_ERR_ARG, "tile width %d not available for n1 = %d", v[1], c->n1);
  if (v[2] > 0 && zpass_blocks(c->prec, c->H, v[2], nl) < 0)
    return fail(XPFC_ERR_ARG, "real-space block %d not available for n2 = %d", v[2], c->n2);
  if (v[0] > 0) c->tc_x = v[0];
  if (c->n1 > 1 && v[1] > 0) c->tc_y = v[1];
  if (v[2] > 0) c->ntz = v[2];
  c->tuned = true;  // explicit choice: the autotuner leaves it alone
  drop_graphs(c);
  return XPFC_OK;
}

extern "C" int xpfc_get_variants(const xpfc_ctx* c, int* v) {
  if (!c || !v) return fail(XPFC_ERR_ARG, "null argument");
  if (!c->cls.empty()) {
    v[0] = v[1] = v[2] = 0;
    return XPFC_OK;
  }
  v[0] = c->tc_x;
  v[1] = c->tc_y;
  v[2] = c->ntz;
  return XPFC_OK;
}

extern "C" int xpfc_set_graphs(xpfc_ctx* c, int on) {
  if (!c) return fail(XPFC_ERR_ARG, "null context");
  c->use_graphs = on != 0;
  if (!c->use_graphs) {
    drop_graphs(c);
    for (auto& g : c->rgraphs) (void)hipGraphExecDestroy(g.second);
    c->rgraphs.clear();
  }
  return XPFC_OK;
}

extern "C" int xpfc_step(xpfc_ctx* c, int nsteps) {
  if (!c) return fail(XPFC_ERR_ARG, "null context");
  if (nsteps < 0) return fail(XPFC_ERR_ARG, "nsteps < 0");
  if (!c->cls.empty()) return mixed_step(c, nsteps);
  if (c->active.empty() || nsteps == 0) return XPFC_OK;
  for (int s : c->active)
    if (!c->inst[s].have_field) return fail(XPFC_ERR_STATE, "instance %d: no field set", s);
  if (!c->tuned && autotune_enabled()) {
    int rc = autotune(c);
    if (rc) return rc;
  }
  if (int rc = initial_records(c)) return rc;
  auto any_samples = [&](long long off) {
    for (int s : c->active)
      if (inst_samples(c->inst[s], c->inst[s].step + off)) return true;
    return false;
  };
  long long left = nsteps;
  while (left > 0) {
    if (any_samples(0)) {
      if ((int)c->pending.size() == SUM_ROWS) {
        int rc = drain(c);
        if (rc) return rc;
      }
      XPending pr;
      pr.row = (int)c->pending.size();
      for (int s : c->active)
        if (inst_samples(c->inst[s], c->inst[s].step)) pr.who.emplace_back(s, c->inst[s].step);
      int rc = launch_one_step(c, true, pr.row);
      if (rc) return rc;
      c->pending.push_back(pr);
      advance(c, 1);
      --left;
      continue;
    }
    // a run of unsampled steps up to the next sampled one
    long long run = 1;
    while (run < left && !any_samples(run)) ++run;
    int rc = run_plain_steps(c, run);
    if (rc) return rc;
    left -= run;
  }
  return XPFC_OK;  // samples are read back by the next call that needs them
}

extern "C" long long xpfc_inst_current_step(const xpfc_ctx* c, int s) {
  if (!c || s < 0 || s >= c->K) return 0;
  if (!c->cls.empty()) return xpfc_inst_current_step(c->cls[c->where[s].first], c->where[s].second);
  return c->inst[s].step;
}
extern "C" long long xpfc_current_step(const xpfc_ctx* c) { return xpfc_inst_current_step(c, 0); }

extern "C" int xpfc_time_steps(xpfc_ctx* c, int nsteps, float* ms) {
  if (!c || !ms) return fail(XPFC_ERR_ARG, "null argument");
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, c->stream));
  int rc = xpfc_step(c, nsteps);
  HIPCHK(hipEventRecord(e1, c->stream));
  HIPCHK(hipEventSynchronize(e1));
  HIPCHK(hipEventElapsedTime(ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  return c->cls.empty() ? drain_report(c) : mixed_sync(c);
}

extern "C" int xpfc_copy_bandwidth(int device, size_t bytes, int reps, double* gbs) {
  if (!gbs || reps < 1 || bytes < 16) return fail(XPFC_ERR_ARG, "bad arguments");
  bytes &= ~(size_t)15;
  if (hipSetDevice(device) != hipSuccess) return fail(XPFC_ERR_DEVICE, "hipSetDevice(%d) failed", device);
  struct Res {
    void *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Res() {
      if (s) (void)hipStreamSynchronize(s);
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      if (a) (void)hipFree(a);
      if (b) (void)hipFree(b);
      if (s) (void)hipStreamDestroy(s);
    }
  } r;
  if (hipMalloc(&r.a, bytes) != hipSuccess || hipMalloc(&r.b, bytes) != hipSuccess)
    return fail(XPFC_ERR_NOMEM, "device allocation of 2 x %zu bytes failed", bytes);
  HIPCHK(hipStreamCreateWithFlags(&r.s, hipStreamNonBlocking));
  HIPCHK(hipEventCreate(&r.e0));
  HIPCHK(hipEventCreate(&r.e1));
  HIPCHK(hipMemsetAsync(r.a, 0, bytes, r.s));
  LCHK(launch_stream_copy(r.a, r.b, bytes, r.s));  // warm-up
  HIPCHK(hipEventRecord(r.e0, r.s));
  for (int i = 0; i < reps; ++i) LCHK(launch_stream_copy(r.a, r.b, bytes, r.s));
  HIPCHK(hipEventRecord(r.e1, r.s));
  HIPCHK(hipEventSynchronize(r.e1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, r.e0, r.e1));
  *gbs = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
  return XPFC_OK;
}

extern "C" int xpfc_time_passes(xpfc_ctx* c, int nsteps, double* ms) {
  XPFC_NOT_MIXED(c, "xpfc_time_passes");
  if (!c || !ms || nsteps <= 0) return fail(XPFC_ERR_ARG, "bad arguments");
  for (int s : c->active)
    if (!c->inst[s].have_field) return fail(XPFC_ERR_STATE, "no field set");
  if (!c->tuned && autotune_enabled()) {
    int rc = autotune(c);
    if (rc) return rc;
  }
  std::vector<hipEvent_t> ev((size_t)nsteps * 4 + 1);
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  HIPCHK(hipEventRecord(ev[0], c->stream));
  for (int i = 0; i < nsteps; ++i) {
    int rc = launch_one_step(c, false, 0, &ev[(size_t)i * 4 + 1]);
    if (rc) return rc;
    advance(c, 1);
  }
  HIPCHK(hipEventSynchronize(ev.back()));
  double acc[4] = {0, 0, 0, 0};
  for (int i = 0; i < nsteps; ++i)
    for (int k = 0; k < 4; ++k) {
      float t = 0.f;
      HIPCHK(hipEventElapsedTime(&t, ev[(size_t)i * 4 + k], ev[(size_t)i * 4 + k + 1]));
      acc[k] += t;
    }
  float tot = 0.f;
  HIPCHK(hipEventElapsedTime(&tot, ev[0], ev.back()));
  for (int k = 0; k < 4; ++k) ms[k] = acc[k] / nsteps;
  ms[4] = (double)tot / nsteps;
  for (auto& e : ev) (void)hipEventDestroy(e);
  return XPFC_OK;
}

extern "C" int xpfc_query(xpfc_ctx* c) {
  if (!c) return -fail(XPFC_ERR_ARG, "null context");
  const hipError_t e = hipStreamQuery(c->stream);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  return -fail(XPFC_ERR_DEVICE, "hipStreamQuery failed: %s", hipGetErrorString(e));
}

extern "C" int xpfc_probe_finite(xpfc_ctx* c, int* finite) {
  if (!c || !finite) return fail(XPFC_ERR_ARG, "null argument");
  if (!c->cls.empty()) return mixed_probe(c, finite);
  LCHK(launch_probe(c->prec, c->psi, c->stride, c->nyoff, c->K, c->d_flags, c->stream));
  HIPCHK(hipMemcpyAsync(finite, c->d_flags, (size_t)c->K * sizeof(int), hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return XPFC_OK;
}

extern "C" int xpfc_sync(xpfc_ctx* c) {
  if (!c) return fail(XPFC_ERR_ARG, "null context");
  if (!c->cls.empty()) return mixed_sync(c);
  HIPCHK(hipStreamSynchronize(c->stream));
  return drain_report(c);
}

// counts on success, minus the status code on failure
extern "C" int xpfc_inst_num_samples(xpfc_ctx* c, int s) {
  XPFC_ROUTE(c, s, xpfc_inst_num_samples(cc, ss));
  if (!c || s < 0 || s >= c->K) return -fail(XPFC_ERR_ARG, "bad arguments");
  int rc = drain(c);
  if (rc) return -rc;
  return (int)c->inst[s].samples.size();
}
extern "C" int xpfc_inst_get_samples(xpfc_ctx* c, int s, xpfc_sample* out, int max) {
  XPFC_ROUTE(c, s, xpfc_inst_get_samples(cc, ss, out, max));
  if (!c || s < 0 || s >= c->K || (max > 0 && !out)) return -fail(XPFC_ERR_ARG, "bad arguments");
  int rc = drain(c);
  if (rc) return -rc;
  const auto& v = c->inst[s].samples;
  const int n = std::min(max, (int)v.size());
  for (int i = 0; i < n; ++i) out[i] = v[i];
  return n;
}
extern "C" void xpfc_inst_clear_samples(xpfc_ctx* c, int s) {
  if (c && !c->cls.empty() && s >= 0 && s < c->K) {
    xpfc_inst_clear_samples(c->cls[c->where[s].first], c->where[s].second);
    return;
  }
  if (c && s >= 0 && s < c->K && drain(c) == XPFC_OK) {
    c->inst[s].samples.clear();
    c->inst[s].nonfinite = c->inst[s].nonfinite_reported = false;
  }
}
extern "C" int xpfc_num_samples(xpfc_ctx* c) { return xpfc_inst_num_samples(c, 0); }
extern "C" int xpfc_get_samples(xpfc_ctx* c, xpfc_sample* out, int max) {
  return xpfc_inst_get_samples(c, 0, out, max);
}
extern "C" void xpfc_clear_samples(xpfc_ctx* c) { xpfc_inst_clear_samples(c, 0); }

// ---------------------------------------------------------------------------
// diagnostics
// ---------------------------------------------------------------------------
extern "C" int xpfc_inst_fe_calc(xpfc_ctx* c, int s, double* F_ref, double* f_mean) {
  XPFC_ROUTE(c, s, xpfc_inst_fe_calc(cc, ss, F_ref, f_mean));
  int rc = check_inst(c, s);
  if (rc) return rc;
  if (!F_ref || !f_mean) return fail(XPFC_ERR_ARG, "null argument");
  if (!c->inst[s].have_field) return fail(XPFC_ERR_STATE, "no field set");
  double m;
  return fe_now(c, s, F_ref, f_mean, &m);
}

extern "C" int xpfc_inst_avg_calc(xpfc_ctx* c, int s, double* mean) {
  XPFC_ROUTE(c, s, xpfc_inst_avg_calc(cc, ss, mean));
  int rc = check_inst(c, s);
  if (rc) return rc;
  if (!mean) return fail(XPFC_ERR_ARG, "null argument");
  if (!c->inst[s].have_field) return fail(XPFC_ERR_STATE, "no field set");
  *mean = c->inst[s].mean;
  return XPFC_OK;
}

extern "C" int xpfc_inst_sse_calc(xpfc_ctx* c, int s, double* rmse) {
  XPFC_ROUTE(c, s, xpfc_inst_sse_calc(cc, ss, rmse));
  int rc = check_inst(c, s);
  if (rc) return rc;
  if (!rmse) return fail(XPFC_ERR_ARG, "null argument");
  if ((rc = drain(c)) != XPFC_OK) return rc;
  if (!c->inst[s].have_rmse) return fail(XPFC_ERR_STATE, "no sampled step yet");
  *rmse = c->inst[s].last_rmse;
  return XPFC_OK;
}

extern "C" int xpfc_fe_calc(xpfc_ctx* c, double* F_ref, double* f_mean) {
  return xpfc_inst_fe_calc(c, 0, F_ref, f_mean);
}
extern "C" int xpfc_avg_calc(xpfc_ctx* c, double* mean) { return xpfc_inst_avg_calc(c, 0, mean); }
extern "C" int xpfc_sse_calc(xpfc_ctx* c, double* rmse) { return xpfc_inst_sse_calc(c, 0, rmse); }

// ---------------------------------------------------------------------------
// operator level (device pointers; instance 0's physics)
// ---------------------------------------------------------------------------
extern "C" int xpfc_malloc(xpfc_ctx* c, void** d, size_t bytes) {
  if (!c || !d) return fail(XPFC_ERR_ARG, "null argument");
  if (hipMalloc(d, bytes) != hipSuccess) return fail(XPFC_ERR_NOMEM, "hipMalloc(%zu) failed", bytes);
  return XPFC_OK;
}
extern "C" int xpfc_free(xpfc_ctx* c, void* d) {
  if (!c) return fail(XPFC_ERR_ARG, "null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipFree(d));
  return XPFC_OK;
}
extern "C" int xpfc_memcpy_h2d(xpfc_ctx* c, void* dst, const void* src, size_t bytes) {
  if (!c) return fail(XPFC_ERR_ARG, "null context");
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return XPFC_OK;
}
extern "C" int xpfc_memcpy_d2h(xpfc_ctx* c, void* dst, const void* src, size_t bytes) {
  if (!c) return fail(XPFC_ERR_ARG, "null context");
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return XPFC_OK;
}

namespace {
int op_done(xpfc_ctx* c) {
  HIPCHK(hipStreamSynchronize(c->stream));
  return XPFC_OK;
}
const Phys& ph0(const xpfc_ctx* c) { return c->inst[0].ia.ph; }
}  // namespace

extern "C" int xpfc_power_functions(xpfc_ctx* c, int p, const void* src, void* dst) {
  XPFC_NOT_MIXED(c, "xpfc_power_functions");
  if (!c) return fail(XPFC_ERR_ARG, "null context");
  if (!src || !dst || p < 0) return fail(XPFC_ERR_ARG, "bad power_functions arguments");
  LCHK(launch_power(c->prec, src, dst, c->N, p, c->inst[0].p.powf_fp32 != 0, c->stream));
  return op_done(c);
}

extern "C" int xpfc_convolve_fields(xpfc_ctx* c, const void* psi_k, void* out) {
  XPFC_NOT_MIXED(c, "xpfc_convolve_fields");
  if (!c) return fail(XPFC_ERR_ARG, "null context");
  if (!psi_k || !out) return fail(XPFC_ERR_ARG, "null buffer");
  LCHK(launch_convolve(c->prec, psi_k, out, c->g, ph0(c), c->n2, c->stream));
  return op_done(c);
}

extern "C" int xpfc_euler_step(xpfc_ctx* c, const void* psi_k, const void* p2_k, const void* p3_k,
                               void* out) {
  XPFC_NOT_MIXED(c, "xpfc_euler_step");
  if (!c) return fail(XPFC_ERR_ARG, "null context");
  if (!psi_k || !p2_k || !p3_k || !out) return fail(XPFC_ERR_ARG, "null buffer");
  const xpfc_params& p = c->inst[0].p;
  LCHK(launch_euler(c->prec, psi_k, p2_k, p3_k, out, c->g, ph0(c), p.a, p.b, c->n2, c->stream));
  return op_done(c);
}

extern "C" int xpfc_residual_calc(xpfc_ctx* c, const void* kp1, const void* k0, const void* p2,
                                  const void* p3, void* out) {
  XPFC_NOT_MIXED(c, "xpfc_residual_calc");
  if (!c) return fail(XPFC_ERR_ARG, "null context");
  if (!kp1 || !k0 || !p2 || !p3 || !out) return fail(XPFC_ERR_ARG, "null buffer");
  const xpfc_params& p = c->inst[0].p;
  LCHK(launch_residual(c->prec, kp1, k0, p2, p3, out, c->g, ph0(c), p.a, p.b, c->n2, c->stream));
  return op_done(c);
}

extern "C" int xpfc_scaling_step(xpfc_ctx* c, void* buf, int spectral) {
  XPFC_NOT_MIXED(c, "xpfc_scaling_step");
  if (!c) return fail(XPFC_ERR_ARG, "null context");
  if (!buf) return fail(XPFC_ERR_ARG, "null buffer");
  const long long n = spectral ? 2LL * c->n0 * c->n1 * (c->H + 1) : c->N;
  LCHK(launch_scale(c->prec, buf, n, 1.0 / (double)c->N, c->stream));
  return op_done(c);
}

namespace {
// scratch table entry K: custom real-space constants for operator calls
int scratch_args(xpfc_ctx* c, double mean, double scale) {
  InstArgs ia = c->inst[0].ia;
  ia.za.mean = mean;
  ia.za.scale = scale;
  HIPCHK(hipMemcpyAsync(c->d_ia + c->K, &ia, sizeof(ia), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return XPFC_OK;
}
}  // namespace

extern "C" int xpfc_fft_r2c(xpfc_ctx* c, const void* real_in, void* spec_out) {
  XPFC_NOT_MIXED(c, "xpfc_fft_r2c");
  if (!c) return fail(XPFC_ERR_ARG, "null context");
  if (!real_in || !spec_out) return fail(XPFC_ERR_ARG, "null buffer");
  const long long nl = (long long)c->n0 * c->n1;
  int rc = scratch_args(c, 0.0, 1.0);  // unnormalised, as cuFFT
  if (rc) return rc;
  const Batch bs{c->d_ids + c->K, c->d_ia, 0, 0, 0};  // no instance offsets
  // the caller's real field is contiguous; the spectrum uses the context layout
  const ZLines zin{c->n1, (long long)c->n1 * c->H, c->xs};
  LCHK(launch_zpass(c->prec, Z_R2C_FLUCT, real_in, nullptr, c->aux, ny_of(c, c->aux), nl, c->H, zin, 256,
                    bs, 1, c->twz, c->twp, nullptr, false, c->any_powf, c->stream));
  if ((rc = forward_xy(c, c->aux, bt_one(c, 0), 1, true))) return rc;
  LCHK(launch_unpack(c->prec, c->aux, spec_out, (int)nl, c->H, c->n1, c->xs, c->stream));
  return op_done(c);
}

extern "C" int xpfc_fft_c2r(xpfc_ctx* c, const void* spec_in, void* real_out) {
  XPFC_NOT_MIXED(c, "xpfc_fft_c2r");
  if (!c) return fail(XPFC_ERR_ARG, "null context");
  if (!spec_in || !real_out) return fail(XPFC_ERR_ARG, "null buffer");
  const long long nl = (long long)c->n0 * c->n1;
  int rc = scratch_args(c, 0.0, 1.0);  // unnormalised, as cuFFT
  if (rc) return rc;
  LCHK(launch_pack(c->prec, spec_in, c->aux, (int)nl, c->H, c->n1, c->xs, c->stream));
  if ((rc = inverse_xy(c, c->aux, bt_one(c, 0), 1))) return rc;
  const Batch bs{c->d_ids + c->K, c->d_ia, 0, 0, 0};
  const ZLines zout{c->n1, c->xs, (long long)c->n1 * c->H};
  LCHK(launch_zpass(c->prec, Z_C2R, c->aux, ny_of(c, c->aux), real_out, nullptr, nl, c->H, zout, 256,
                    bs, 1, c->twz, c->twp, nullptr, false, c->any_powf, c->stream));
  return op_done(c);
}

// ---------------------------------------------------------------------------
// VTK output (write_vtk.cu:4-35)
// ---------------------------------------------------------------------------
extern "C" int xpfc_inst_write_vtk(xpfc_ctx* c, int s, const char* path) {
  XPFC_ROUTE(c, s, xpfc_inst_write_vtk(cc, ss, path));
  int rc = check_inst(c, s);
  if (rc) return rc;
  if (!path) return fail(XPFC_ERR_ARG, "null path");
  std::vector<double> psi((size_t)c->N);
  if ((rc = xpfc_inst_get_field(c, s, psi.data()))) return rc;
  std::ofstream f(path);
  if (!f) return fail(XPFC_ERR_IO, "cannot open %s for writing", path);
  const xpfc_params& p = c->inst[s].p;
  const int nx = p.n[0], ny = p.n[1], nz = p.dim == 3 ? p.n[2] : 1;
  // write_vtk.cu:15-27 (default ostream formatting, x fastest)
  f << "# vtk DataFile Version 3.0\n";
  f << "Density Data\n";
  f << "ASCII\n";
  f << "DATASET STRUCTURED_POINTS\n";
  f << "DIMENSIONS " << nx << " " << ny << " " << nz << "\n";
  f << "ORIGIN 0 0 0\n";
  f << "SPACING " << double(p.L[0]) / double(nx) << " " << double(p.L[1]) / double(ny) << " "
    << (p.dim == 3 ? double(p.L[2]) / double(nz) : 1.0) << "\n";
  f << "POINT_DATA " << (long long)nx * ny * nz << "\n";
  f << "SCALARS DensityData float\n";
  f << "LOOKUP_TABLE default\n";
  for (int l = 0; l < nz; ++l)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) f << psi[((size_t)i * ny + j) * nz + l] << "\n";
  if (!f) return fail(XPFC_ERR_IO, "write to %s failed", path);
  return XPFC_OK;
}

extern "C" int xpfc_write_vtk(xpfc_ctx* c, const char* path) { return xpfc_inst_write_vtk(c, 0, path); }

// ---------------------------------------------------------------------------
// binary checkpoints: header, then the state (in the context's form) and the
// work spectrum of the instance exactly as they sit in device memory (packed
// half-spectrum layout)
// ---------------------------------------------------------------------------
namespace {
struct CkptHeader {
  char magic[8];          // "XPFCCKP1"
  int32_t version;        // 3
  int32_t precision;
  int32_t n0, n1, n2;     // packed-layout extents
  int32_t state_form;     // KGeom::xreal of the writing context
  int64_t step;
  double mean;
  int64_t stride;         // complex elements per spectrum
  int64_t csize;          // bytes per complex element
  xpfc_params params;
};
const char kMagic[8] = {'X', 'P', 'F', 'C', 'C', 'K', 'P', '1'};
}  // namespace

extern "C" int xpfc_inst_save_checkpoint(xpfc_ctx* c, int s, const char* path) {
  XPFC_ROUTE(c, s, xpfc_inst_save_checkpoint(cc, ss, path));
  int rc = check_inst(c, s);
  if (rc) return rc;
  if (!path) return fail(XPFC_ERR_ARG, "null path");
  if (!c->inst[s].have_field) return fail(XPFC_ERR_STATE, "no field set");
  const XInst& in = c->inst[s];
  CkptHeader h{};
  std::memcpy(h.magic, kMagic, 8);
  h.version = 3;  // 2: xpfc_params with powf_fp32; 3: the state form
  h.state_form = c->g.xreal;
  h.precision = c->prec;
  h.n0 = c->n0;
  h.n1 = c->n1;
  h.n2 = c->n2;
  h.step = in.step;
  h.mean = in.mean;
  h.stride = c->stride;
  h.csize = (int64_t)c->csize;
  h.params = in.p;
  const size_t bytes = (size_t)c->stride * c->csize;
  std::vector<char> buf(2 * bytes);
  HIPCHK(hipMemcpyAsync(buf.data(), inst_ptr(c, c->psi, s), bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(buf.data() + bytes, inst_ptr(c, c->work, s), bytes, hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(XPFC_ERR_IO, "cannot open %s for writing", path);
  const bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1 && std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
  if (std::fclose(f) != 0 || !ok) return fail(XPFC_ERR_IO, "write to %s failed", path);
  return XPFC_OK;
}

extern "C" int xpfc_inst_load_checkpoint(xpfc_ctx* c, int s, const char* path) {
  XPFC_ROUTE(c, s, xpfc_inst_load_checkpoint(cc, ss, path));
  int rc = check_inst(c, s);
  if (rc) return rc;
  if (!path) return fail(XPFC_ERR_ARG, "null path");
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(XPFC_ERR_IO, "cannot open %s", path);
  CkptHeader h{};
  const size_t bytes = (size_t)c->stride * c->csize;
  std::vector<char> buf(2 * bytes);
  bool ok = std::fread(&h, sizeof(h), 1, f) == 1;
  ok = ok && std::memcmp(h.magic, kMagic, 8) == 0 && h.version == 3;
  if (!ok) {
    std::fclose(f);
    return fail(XPFC_ERR_PARSE, "%s: not an XPFC checkpoint", path);
  }
  if (h.params.dim == 2) {  // as xpfc_inst_set_params normalises 2-D params
    h.params.n[2] = 1;
    h.params.L[2] = 1.0;
  }
  if (h.precision != c->prec || h.n0 != c->n0 || h.n1 != c->n1 || h.n2 != c->n2 ||
      h.stride != c->stride || h.csize != (int64_t)c->csize || !same_grid(c->p, h.params)) {
    std::fclose(f);
    return fail(XPFC_ERR_ARG, "%s: grid, box or precision differ from the context's", path);
  }
  if (h.state_form != 0 && h.state_form != 1) {
    std::fclose(f);
    return fail(XPFC_ERR_PARSE, "%s: unknown state form %d", path, (int)h.state_form);
  }
  ok = std::fread(buf.data(), 1, buf.size(), f) == buf.size();
  std::fclose(f);
  if (!ok) return fail(XPFC_ERR_PARSE, "%s: truncated", path);
  if ((rc = drain(c)) != XPFC_OK) return rc;
  XInst& in = c->inst[s];
  in.mean = h.mean;
  if ((rc = build_inst(c, h.params, &in)) != XPFC_OK) return rc;
  if ((rc = upload_inst(c, s)) != XPFC_OK) return rc;
  HIPCHK(hipMemcpyAsync(inst_ptr(c, c->psi, s), buf.data(), bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(inst_ptr(c, c->work, s), buf.data() + bytes, bytes, hipMemcpyHostToDevice,
                        c->stream));
  if (h.state_form != c->g.xreal) {
    // written in the other state form (the working buffer has one form in
    // both): psi_k -> state = unnormalised inverse FFT along axis 0, or
    // state -> psi_k = FFT_0(state) / n0 with the k = 0 slot exactly 0, as the
    // axis-0 pass forms it
    const Batch bt = bt_one(c, s);
    const int dir = c->g.xreal ? +1 : -1;
    LCHK(launch_strided(c->prec, c->psi, x_segs(c, c->tc_x), c->n0, c->tc_x, dir, c->tw0, bt, 1, false,
                        c->stream));
    if (!c->g.xreal) {
      void* ps = inst_ptr(c, c->psi, s);
      LCHK(launch_scale(c->prec, ps, 2LL * c->stride, c->g.inv_n0, c->stream));
      HIPCHK(hipMemsetAsync(ps, 0, c->csize, c->stream));
    }
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  in.step = h.step;
  in.samples.clear();
  in.have_rmse = false;
  in.have_field = true;
  return XPFC_OK;
}

extern "C" int xpfc_save_checkpoint(xpfc_ctx* c, const char* path) {
  return xpfc_inst_save_checkpoint(c, 0, path);
}
extern "C" int xpfc_load_checkpoint(xpfc_ctx* c, const char* path) {
  return xpfc_inst_load_checkpoint(c, 0, path);
}

// ---------------------------------------------------------------------------
// contexts of several grids (xpfc_create_mixed): the instances of each grid
// form a class context; all classes share one stream and advance together
// with one ragged launch per pass (every class's tiles in one grid of
// workgroups), captured in hipGraph step blocks like the one-grid path
// ---------------------------------------------------------------------------
namespace {

// (dim, extents, box) of a parameter set, 2-D normalised
std::vector<double> grid_key(const xpfc_params& p0) {
  xpfc_params p = p0;
  if (p.dim == 2) {
    p.n[2] = 1;
    p.L[2] = 1.0;
  }
  return {(double)p.dim, (double)p.n[0], (double)p.n[1], (double)p.n[2], p.L[0], p.L[1], p.L[2],
          (double)p.precision, (double)p.device};
}

std::vector<int> active_counts(const xpfc_ctx* c) {
  std::vector<int> k;
  for (const xpfc_ctx* q : c->cls) k.push_back((int)q->active.size());
  return k;
}

// segments with the tile counts of tile width tc (as the launchers fill them)
Segs filled(const Segs& in, int tc) {
  Segs s = in;
  int tiles[2];
  for (int i = 0; i < 2; ++i) {
    s.s[i].ntc = (s.s[i].ncols + tc - 1) / tc;
    tiles[i] = (s.s[i].ncols <= 0 || s.s[i].nbatch <= 0) ? 0 : s.s[i].ntc * s.s[i].nbatch;
  }
  s.tiles0 = tiles[0];
  s.tiles1 = tiles[1];
  s.lgc = -1;
  return s;
}

// Workgroup order of a class's strided-pass segment of n workgroups (tiles of
// all its instances, instance-major): runs of up to 2^XCD_RUN_X neighbouring
// tiles per XCD, at least two rounds of them (as the one-grid axis-0 pass).
int rag_lgc(int n) {
  int lg = XCD_RUN_X;
  while (lg >= 0 && n < (16 << lg)) --lg;
  return lg >= 2 ? lg : -1;
}
int round8(int n) { return (n + 7) & ~7; }

// Groups the classes of one pass into launches by (line family, LUT form),
// numbering each launch's workgroups from 0 (wg0 of the strided passes on a
// multiple of 8, so a workgroup's XCD is its index in the class mod 8).
struct RagEntry {
  int grp;   // stream group
  RClass r;  // the class (or half of its active list)
};

void rag_group(std::vector<RagEntry>& in, bool strided, std::vector<RClass>& all,
               std::vector<xpfc_ctx::RagLaunch>& out, int prec) {
  for (int grp : {0, 1})
  for (int fam : {1, 3, 5}) {
    for (int lut : {1, 0}) {
      xpfc_ctx::RagLaunch l;
      l.grp = grp;
      l.fam = fam;
      l.lut = lut != 0;
      l.first = (int)all.size();
      for (auto& e : in) {
        RClass r = e.r;
        if (e.grp != grp || line_family(r.L) != fam || (r.cubic != 0) != l.lut) continue;
        r.wg0 = strided ? round8(l.nwg) : l.nwg;
        l.nwg = r.wg0 + r.nact * r.wgs;
        l.lds = std::max(l.lds, strided ? rag_strided_lds(prec, r.L) : rag_zpass_lds(prec, r.L));
        all.push_back(r);
        ++l.ncls;
      }
      if (l.ncls) out.push_back(l);
    }
  }
}

// The class tables of the current active lists (built once per distinct
// list lengths; the kernels read the lists themselves from device memory).
int mixed_tables(xpfc_ctx* c, bool budget, const xpfc_ctx::RagTables** out) {
  std::vector<int> key = active_counts(c);
  key.push_back(budget ? 1 : 0);
  auto it = c->rtab.find(key);
  if (it != c->rtab.end()) {
    *out = &it->second;
    return XPFC_OK;
  }
  xpfc_ctx::RagTables t;
  std::vector<RagEntry> X, Y, Z;  // the entries of each pass
  const int nc = (int)c->cls.size();
  t.nbx.assign(nc, 0);
  t.nbz.assign(nc, 0);
  // Each class's active list is split in halves between the two stream
  // groups (a single instance goes to the group with less work so far).
  const bool two = c->mixed_groups == 2 && c->stream2;
  double work[2] = {0.0, 0.0};
  // the entries of one pass for a class: one per group that takes a share
  // (h[g] instances, the group-0 share first in the class's active list)
  auto add = [&](std::vector<RagEntry>& list, const RClass& base, const int* h, bool strided_lgc) {
    int first = 0;
    for (int g = 0; g < 2; ++g) {
      if (h[g] == 0) continue;
      RClass e = base;
      e.nact = h[g];
      e.bt.slots = base.bt.slots + first;
      if (strided_lgc) e.lgc = rag_lgc(h[g] * base.wgs);
      list.push_back({g, e});
      first += h[g];
    }
  };
  for (int k = 0; k < nc; ++k) {
    xpfc_ctx* q = c->cls[k];
    const int na = (int)q->active.size();
    t.nbx[k] = rag_strided_wgs(q->prec, q->n0, q->segX);
    t.nbz[k] = rag_zpass_wgs(q->prec, q->H, (long long)q->n0 * q->n1);
    const int nby = q->n1 > 1 ? rag_strided_wgs(q->prec, q->n1, q->segY) : 0;
    if (t.nbx[k] <= 0 || t.nbz[k] <= 0 || nby < 0)
      return fail(XPFC_ERR_ARG, "grid %d x %d x %d has no one-launch form", q->n0, q->n1, q->n2);
    if (na == 0) continue;
    const Batch bt = bt_active(q);
    // the class's share per group
    int h[2] = {na, 0};
    const double per = (double)q->stride * q->csize;
    if (two) {
      const int lo = work[0] <= work[1] ? 0 : 1;
      h[lo] = (na + 1) / 2;
      h[1 - lo] = na - h[lo];
    }
    work[0] += per * h[0];
    work[1] += per * h[1];
    RClass r{};
    r.L = q->n0;
    r.lgc = rag_lgc(na * t.nbx[k]);
    r.wgs = t.nbx[k];
    r.nact = na;
    r.cubic = q->g.cubic;
    r.segs = filled(q->segX, rag_tile(q->n0));
    r.g = q->g;
    r.bt = bt;
    r.a = q->work;
    r.b = q->psi;
    r.tw = q->tw0;
    r.partial = q->px;
    r.budget = budget ? q->d_budget : nullptr;
    add(X, r, h, true);
    if (q->n1 > 1) {
      RClass y{};
      y.L = q->n1;
      y.lgc = -1;
      y.wgs = nby;
      y.nact = na;
      y.segs = filled(q->segY, rag_tile(q->n1));
      y.bt = bt;
      y.a = q->work;
      y.b = q->g.xreal ? q->psi : q->work;  // the inverse pass reads the state in the axis-0 form
      y.tw = q->tw1;
      y.budget = r.budget;
      add(Y, y, h, false);
    }
    RClass z{};
    z.L = q->H;
    z.wgs = t.nbz[k];
    z.nact = na;
    z.pw = q->any_powf;
    z.bt = bt;
    z.a = q->work;
    z.b = (q->n1 > 1 || !q->g.xreal) ? q->work : q->psi;  // 2-D axis-0 form: reads the state
    z.tw = q->twz;
    z.twp = q->twp;
    z.partial = q->pz;
    z.nlines = (long long)q->n0 * q->n1;
    z.zl = ZLines{q->n1, q->xs, q->xs};
    z.nyoff = q->nyoff;
    z.budget = r.budget;
    add(Z, z, h, false);
  }
  if ((int)Z.size() > 2 * RAG_MAX) return fail(XPFC_ERR_ARG, "more than %d grids active at once", RAG_MAX);
  t.two = work[1] > 0.0;
  std::vector<RClass> all;
  rag_group(X, true, all, t.x, c->prec);
  rag_group(Y, true, all, t.y, c->prec);
  rag_group(Z, false, all, t.z, c->prec);
  for (int fam : {1, 3, 5}) LCHK(rag_prepare(c->prec, fam));
  if (!all.empty()) {
    if (hipMalloc((void**)&t.d, all.size() * sizeof(RClass)) != hipSuccess)
      return fail(XPFC_ERR_NOMEM, "class table allocation failed");
    if (hipMemcpy(t.d, all.data(), all.size() * sizeof(RClass), hipMemcpyHostToDevice) != hipSuccess) {
      (void)hipFree(t.d);
      return fail(XPFC_ERR_DEVICE, "class table upload failed");
    }
  }
  *out = &(c->rtab[key] = t);
  return XPFC_OK;
}

bool mixed_pw(const xpfc_ctx* c) {
  for (const xpfc_ctx* q : c->cls)
    if (!q->active.empty() && q->any_powf) return true;
  return false;
}

// The launches of one step of every active instance of every class, one per
// line family present in the pass: axis 0 (separate launches for square /
// cubic and other boxes), axis 1 inverse (3-D classes), real space, axis 1
// forward (3-D classes), then for sampled steps each listed class's finalize
// into its row `rows[k]`.
// The second group's work runs on stream2, forked from and joined back into
// the context stream (also inside a capture, where the two become branches of
// the graph).
int mixed_fork(xpfc_ctx* c, const xpfc_ctx::RagTables& t) {
  if (!t.two) return XPFC_OK;
  HIPCHK(hipEventRecord(c->ev_fork, c->stream));
  HIPCHK(hipStreamWaitEvent(c->stream2, c->ev_fork, 0));
  return XPFC_OK;
}
int mixed_join(xpfc_ctx* c, const xpfc_ctx::RagTables& t) {
  if (!t.two) return XPFC_OK;
  HIPCHK(hipEventRecord(c->ev_join, c->stream2));
  HIPCHK(hipStreamWaitEvent(c->stream, c->ev_join, 0));
  return XPFC_OK;
}

int mixed_one_step(xpfc_ctx* c, const xpfc_ctx::RagTables& t, bool sample, int sib,
                   const std::vector<int>* rows, bool pw, bool forkjoin = true, int only = -1) {
  if (forkjoin)
    if (int rc = mixed_fork(c, t)) return rc;
  auto st = [&](const xpfc_ctx::RagLaunch& l) { return l.grp ? c->stream2 : c->stream; };
  auto mine = [&](const xpfc_ctx::RagLaunch& l) { return only < 0 || l.grp == only; };
  for (const auto& l : t.x)
    if (mine(l))
      LCHK(launch_xstep_ragged(c->prec, l.fam, t.d + l.first, l.ncls, l.nwg, l.lds, l.lut, sample, sib, st(l)));
  for (const auto& l : t.y)
    if (mine(l)) LCHK(launch_strided_ragged(c->prec, l.fam, t.d + l.first, l.ncls, l.nwg, l.lds, +1, sib, st(l)));
  for (const auto& l : t.z)
    if (mine(l))
      LCHK(launch_zpass_ragged(c->prec, l.fam, t.d + l.first, l.ncls, l.nwg, l.lds, sample, pw, sib, st(l)));
  for (const auto& l : t.y)
    if (mine(l)) LCHK(launch_strided_ragged(c->prec, l.fam, t.d + l.first, l.ncls, l.nwg, l.lds, -1, sib, st(l)));
  if (forkjoin)
    if (int rc = mixed_join(c, t)) return rc;
  if (sample && rows) {
    for (size_t k = 0; k < c->cls.size(); ++k) {
      if ((*rows)[k] < 0) continue;
      xpfc_ctx* q = c->cls[k];
      LCHK(launch_finalize(q->px, t.nbx[k], q->pz, t.nbz[k], bt_active(q), (int)q->active.size(),
                           q->sums + (size_t)(*rows)[k] * q->K * NSUMS, c->stream));
    }
  }
  return XPFC_OK;
}

void mixed_drop_graphs(xpfc_ctx* c) {
  for (auto& g : c->rgraphs) (void)hipGraphExecDestroy(g.second);
  c->rgraphs.clear();
}

// Unsampled steps: replays of captured blocks of GRAPH_BLOCK steps (or a
// power of two below), keyed by length, active counts and the power form.
int mixed_plain(xpfc_ctx* c, const xpfc_ctx::RagTables& t, long long run, bool pw) {
  while (run > 0) {
    if (!c->use_graphs) {
      int rc = mixed_one_step(c, t, false, 0, nullptr, pw);
      if (rc) return rc;
      for (xpfc_ctx* q : c->cls) advance(q, 1);
      --run;
      continue;
    }
    int len = GRAPH_BLOCK;
    if (run < GRAPH_BLOCK) {
      len = 1;
      while (2 * len <= run) len *= 2;
    }
    // one graph per stream group, captured on that group's stream and
    // launched there between a fork and a join
    hipGraphExec_t ex[2] = {nullptr, nullptr};
    for (int g = 0; g < (t.two ? 2 : 1); ++g) {
      std::vector<int> key = active_counts(c);
      key.push_back(len);
      key.push_back(pw ? 1 : 0);
      key.push_back(t.two ? 10 + g : 0);
      auto it = c->rgraphs.find(key);
      if (it != c->rgraphs.end()) {
        ex[g] = it->second;
        continue;
      }
      if (c->rgraphs.size() >= 64) {
        mixed_drop_graphs(c);
        ex[0] = nullptr;
        g = -1;  // rebuild both
        continue;
      }
      hipStream_t sg = g ? c->stream2 : c->stream;
      hipGraph_t graph = nullptr;
      HIPCHK(hipStreamBeginCapture(sg, hipStreamCaptureModeThreadLocal));
      int rc = XPFC_OK;
      for (int i = 0; i < len && rc == XPFC_OK; ++i)
        rc = mixed_one_step(c, t, false, i, nullptr, pw, false, t.two ? g : -1);
      const hipError_t ec = hipStreamEndCapture(sg, &graph);
      if (rc != XPFC_OK || ec != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc != XPFC_OK ? rc : fail(XPFC_ERR_DEVICE, "step capture failed: %s", hipGetErrorString(ec));
      }
      const hipError_t ei = hipGraphInstantiate(&ex[g], graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ei != hipSuccess) return fail(XPFC_ERR_DEVICE, "graph instantiation failed: %s", hipGetErrorString(ei));
      c->rgraphs[key] = ex[g];
    }
    if (int rc = mixed_fork(c, t)) return rc;
    HIPCHK(hipGraphLaunch(ex[0], c->stream));
    if (t.two) HIPCHK(hipGraphLaunch(ex[1], c->stream2));
    if (int rc = mixed_join(c, t)) return rc;
    for (xpfc_ctx* q : c->cls) advance(q, len);
    run -= len;
  }
  return XPFC_OK;
}

// One block of `len` steps in which slot i of class k takes hbud[i] of them
// (the kernels of steps past an instance's budget return at once): the
// budgets go to the device, then the block runs as one captured graph.
int mixed_budget_block(xpfc_ctx* c, int len) {
  const xpfc_ctx::RagTables* t = nullptr;
  if (int rc = mixed_tables(c, true, &t)) return rc;
  const bool pw = mixed_pw(c);
  for (xpfc_ctx* q : c->cls)
    HIPCHK(hipMemcpyAsync(q->d_budget, q->hbud.data(), q->hbud.size() * sizeof(int), hipMemcpyHostToDevice,
                          c->stream));
  if (!c->use_graphs) {
    if (int rc = mixed_fork(c, *t)) return rc;
    for (int i = 0; i < len; ++i)
      if (int rc = mixed_one_step(c, *t, false, i, nullptr, pw, false)) return rc;
    if (int rc = mixed_join(c, *t)) return rc;
  } else {
    std::vector<int> key = active_counts(c);
    key.push_back(len);
    key.push_back(pw ? 1 : 0);
    key.push_back(-1);  // budgeted block
    hipGraphExec_t ex = nullptr;
    auto it = c->rgraphs.find(key);
    if (it != c->rgraphs.end()) {
      ex = it->second;
    } else {
      if (c->rgraphs.size() >= 64) mixed_drop_graphs(c);
      hipGraph_t graph = nullptr;
      HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      int rc = mixed_fork(c, *t);
      for (int i = 0; i < len && rc == XPFC_OK; ++i) rc = mixed_one_step(c, *t, false, i, nullptr, pw, false);
      if (rc == XPFC_OK) rc = mixed_join(c, *t);
      const hipError_t ec = hipStreamEndCapture(c->stream, &graph);
      if (rc != XPFC_OK || ec != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc != XPFC_OK ? rc : fail(XPFC_ERR_DEVICE, "step capture failed: %s", hipGetErrorString(ec));
      }
      const hipError_t ei = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ei != hipSuccess) return fail(XPFC_ERR_DEVICE, "graph instantiation failed: %s", hipGetErrorString(ei));
      c->rgraphs[key] = ex;
    }
    HIPCHK(hipGraphLaunch(ex, c->stream));
  }
  for (xpfc_ctx* q : c->cls)
    for (int s : q->active) q->inst[s].step += q->hbud[s];
  return XPFC_OK;
}

int mixed_step(xpfc_ctx* c, int nsteps) {
  if (nsteps == 0) return XPFC_OK;
  bool any = false;
  for (xpfc_ctx* q : c->cls) {
    for (int s : q->active)
      if (!q->inst[s].have_field) return fail(XPFC_ERR_STATE, "an instance of grid %d x %d x %d has no field",
                                              q->n0, q->n1, q->n2);
    any = any || !q->active.empty();
  }
  if (!any) return XPFC_OK;
  for (xpfc_ctx* q : c->cls)
    if (int rc = initial_records(q)) return rc;
  const xpfc_ctx::RagTables* t = nullptr;
  if (int rc = mixed_tables(c, false, &t)) return rc;
  const bool pw = mixed_pw(c);
  auto samples_at = [&](long long off) {
    for (const xpfc_ctx* q : c->cls)
      for (int s : q->active)
        if (inst_samples(q->inst[s], q->inst[s].step + off)) return true;
    return false;
  };
  long long left = nsteps;
  while (left > 0) {
    if (samples_at(0)) {
      std::vector<int> rows(c->cls.size(), -1);
      for (size_t k = 0; k < c->cls.size(); ++k) {
        xpfc_ctx* q = c->cls[k];
        XPending pr;
        for (int s : q->active)
          if (inst_samples(q->inst[s], q->inst[s].step)) pr.who.emplace_back(s, q->inst[s].step);
        if (pr.who.empty()) continue;
        if ((int)q->pending.size() == SUM_ROWS)
          if (int rc = drain(q)) return rc;
        pr.row = (int)q->pending.size();
        rows[k] = pr.row;
        q->pending.push_back(pr);
      }
      if (int rc = mixed_one_step(c, *t, true, 0, &rows, pw)) return rc;
      for (xpfc_ctx* q : c->cls) advance(q, 1);
      --left;
      continue;
    }
    long long run = 1;
    while (run < left && !samples_at(run)) ++run;
    if (int rc = mixed_plain(c, *t, run, pw)) return rc;
    left -= run;
  }
  return XPFC_OK;
}

int mixed_set_active(xpfc_ctx* c, const int* slots, int n) {
  std::vector<std::vector<int>> per(c->cls.size());
  if (!slots) {
    for (int s = 0; s < c->K; ++s) per[c->where[s].first].push_back(c->where[s].second);
  } else {
    if (n < 0 || n > c->K) return fail(XPFC_ERR_ARG, "bad active count");
    std::vector<char> seen(c->K, 0);
    for (int i = 0; i < n; ++i) {
      if (slots[i] < 0 || slots[i] >= c->K || seen[slots[i]])
        return fail(XPFC_ERR_ARG, "bad or repeated instance %d", slots[i]);
      seen[slots[i]] = 1;
      per[c->where[slots[i]].first].push_back(c->where[slots[i]].second);
    }
  }
  const int none = 0;
  for (size_t k = 0; k < c->cls.size(); ++k)
    if (int rc = xpfc_set_active(c->cls[k], per[k].empty() ? &none : per[k].data(), (int)per[k].size()))
      return rc;
  return XPFC_OK;
}

int mixed_sync(xpfc_ctx* c) {
  HIPCHK(hipStreamSynchronize(c->stream));
  int worst = XPFC_OK;
  for (xpfc_ctx* q : c->cls) {
    const int rc = drain_report(q);
    if (rc == XPFC_ERR_NONFINITE) worst = rc;  // keep reading the other classes
    else if (rc) return rc;
  }
  return worst;
}

int mixed_probe(xpfc_ctx* c, int* finite) {
  for (size_t k = 0; k < c->cls.size(); ++k) {
    xpfc_ctx* q = c->cls[k];
    std::vector<int> f((size_t)q->K);
    if (int rc = xpfc_probe_finite(q, f.data())) return rc;
    for (int s = 0; s < c->K; ++s)
      if (c->where[s].first == (int)k) finite[s] = f[(size_t)c->where[s].second];
  }
  return XPFC_OK;
}

}  // namespace

extern "C" int xpfc_create_mixed(const xpfc_params* ps, int K, xpfc_ctx** out) {
  if (!ps || !out || K < 1) return fail(XPFC_ERR_ARG, "bad arguments");
  *out = nullptr;
  for (int s = 1; s < K; ++s)
    if (ps[s].precision != ps[0].precision || ps[s].device != ps[0].device)
      return fail(XPFC_ERR_ARG, "instance %d: precision or device differ from instance 0", s);
  // classes in order of first appearance
  std::vector<std::vector<double>> keys;
  std::vector<std::vector<int>> members;
  for (int s = 0; s < K; ++s) {
    const std::vector<double> k = grid_key(ps[s]);
    size_t i = 0;
    while (i < keys.size() && keys[i] != k) ++i;
    if (i == keys.size()) {
      keys.push_back(k);
      members.emplace_back();
    }
    members[i].push_back(s);
  }
  if ((int)keys.size() > RAG_MAX) return fail(XPFC_ERR_ARG, "more than %d grids in one context", RAG_MAX);
  xpfc_ctx* c = new (std::nothrow) xpfc_ctx();
  if (!c) return fail(XPFC_ERR_NOMEM, "host allocation failed");
  auto bail = [&](int r) {
    free_all(c);
    delete c;
    return r;
  };
  c->p = ps[0];
  c->K = K;
  c->prec = ps[0].precision == XPFC_FP32_MIXED ? 1 : 0;
  if (hipSetDevice(ps[0].device) != hipSuccess)
    return bail(fail(XPFC_ERR_DEVICE, "hipSetDevice(%d) failed", ps[0].device));
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(XPFC_ERR_DEVICE, "stream creation failed"));
  // the second stream group (default priority, like the first); XPFC_MIXED_STREAMS=1 keeps one
  const char* ms = std::getenv("XPFC_MIXED_STREAMS");
  c->mixed_groups = (ms && ms[0] == '1' && ms[1] == 0) ? 1 : 2;
  if (c->mixed_groups == 2 &&
      (hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess ||
       hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
       hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess))
    return bail(fail(XPFC_ERR_DEVICE, "stream creation failed"));
  c->where.assign(K, {0, 0});
  for (size_t k = 0; k < members.size(); ++k) {
    std::vector<xpfc_params> gp;
    for (int s : members[k]) gp.push_back(ps[s]);
    xpfc_ctx* q = nullptr;
    if (int rc = create_impl(gp.data(), (int)gp.size(), c->stream, &q)) return bail(rc);
    q->hbud.assign((size_t)q->K, 0);
    c->cls.push_back(q);
    // the ragged kernels compile lines of 8 .. 1024 points (half lines up to
    // 512) of each family
    if (rag_strided_wgs(q->prec, q->n0, q->segX) <= 0 || rag_zpass_wgs(q->prec, q->H, 1) <= 0 ||
        (q->n1 > 1 && rag_strided_wgs(q->prec, q->n1, q->segY) <= 0))
      return bail(fail(XPFC_ERR_ARG, "grid %d x %d x %d has no one-launch form (mixed-grid contexts take "
                       "extents of 2^k, 3*2^k or 5*2^k points up to 1024)", q->n0, q->n1, q->n2));
    c->inst_of.push_back(members[k]);
    for (size_t i = 0; i < members[k].size(); ++i) c->where[members[k][i]] = {(int)k, (int)i};
  }
  const char* eg = std::getenv("XPFC_GRAPHS");
  c->use_graphs = !(eg && eg[0] == '0');
  *out = c;
  return XPFC_OK;
}

extern "C" int xpfc_step_budget(xpfc_ctx* c, int len, const int* steps) {
  if (!c || !steps || len < 1) return fail(XPFC_ERR_ARG, "bad arguments");
  if (c->cls.empty()) {  // one grid
    bool any = false;
    std::fill(c->hbud.begin(), c->hbud.end(), 0);
    for (int s : c->active) {
      const int b = steps[s];
      const XInst& in = c->inst[(size_t)s];
      if (b < 0 || b > len) return fail(XPFC_ERR_ARG, "instance %d: %d steps in a block of %d", s, b, len);
      if (!in.have_field) return fail(XPFC_ERR_STATE, "instance %d: no field set", s);
      for (long long n = in.step; n < in.step + b; ++n)
        if (inst_samples(in, n))
          return fail(XPFC_ERR_ARG, "instance %d samples at step %lld inside the block (use xpfc_step)", s, n);
      c->hbud[(size_t)s] = b;
      any = any || b > 0;
    }
    if (!any) return XPFC_OK;
    if (!c->tuned && autotune_enabled()) {
      if (int rc = autotune(c)) return rc;
    }
    return uniform_budget_block(c, len);
  }
  bool any = false;
  for (size_t k = 0; k < c->cls.size(); ++k) {
    xpfc_ctx* q = c->cls[k];
    std::fill(q->hbud.begin(), q->hbud.end(), 0);
    for (int slot : q->active) {
      const int s = c->inst_of[k][(size_t)slot];
      const int b = steps[s];
      const XInst& in = q->inst[slot];
      if (b < 0 || b > len) return fail(XPFC_ERR_ARG, "instance %d: %d steps in a block of %d", s, b, len);
      if (!in.have_field) return fail(XPFC_ERR_STATE, "instance %d: no field set", s);
      for (long long n = in.step; n < in.step + b; ++n)
        if (inst_samples(in, n))
          return fail(XPFC_ERR_ARG, "instance %d samples at step %lld inside the block (use xpfc_step)", s, n);
      q->hbud[(size_t)slot] = b;
      any = any || b > 0;
    }
  }
  if (!any) return XPFC_OK;
  return mixed_budget_block(c, len);
}

