// Host runtime of libxpfc: parameters, context, initial conditions, the step
// driver, diagnostics, VTK I/O.  Everything device-side goes through the
// launch_* entry points of xpfc_kernels.hip.
#include "../../include/xpfc.h"
#include "xpfc_internal.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <map>
#include <mutex>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

using namespace xpfc;

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

#define HIPCHK(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail(XPFC_ERR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));   \
  } while (0)

#define LCHK(expr)                                                                    \
  do {                                                                                \
    int r_ = (expr);                                                                  \
    if (r_ < 0) {                                                                     \
      hipError_t e_ = hipGetLastError();                                              \
      return fail(XPFC_ERR_DEVICE, "%s failed (%d): %s", #expr, r_,                   \
                  hipGetErrorString(e_));                                             \
    }                                                                                 \
  } while (0)

bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

constexpr int SUM_SLOTS = 1024;  // sampled-step sums held on the device per drain

}  // namespace

struct xpfc_ctx {
  xpfc_params p{};
  int prec = 0;
  int n0 = 0, n1 = 0, n2 = 0, H = 0;
  long long N = 0, nmain = 0, nny = 0;
  size_t csize = 0, rsize = 0;
  void* psi = nullptr;   // packed spectrum of psi - mean (state)
  void* work = nullptr;  // nonlinear term / inverse-transform work field
  void* aux = nullptr;   // diagnostics / field I/O scratch
  void* tw0 = nullptr;   // exp(-2 pi i m / n0)
  void* tw1 = nullptr;   // exp(-2 pi i m / n1)
  void* twz = nullptr;   // exp(-2 pi i m / H)
  void* twp = nullptr;   // exp(-2 pi i k / n2), k < H
  double* px = nullptr;  // X-pass partials
  double* pz = nullptr;  // Z-pass partials
  double* sums = nullptr;  // [SUM_SLOTS][NSUMS]
  int npx_cap = 0, npz_cap = 0;
  int tc_x = 0, tc_y = 0, ntz = 256;  // pass variants (tile width, real-space block)
  bool tuned = false;
  hipStream_t stream = nullptr;
  double mean = 0.0;
  KGeom g{};
  Phys ph{};
  ZArgs za{};
  Segs segX{}, segY{};
  long long step = 0;
  bool have_field = false;
  std::vector<xpfc_sample> samples;
  // pending sampled steps: (slot, step n)
  std::vector<std::pair<int, long long>> pending;
  double last_rmse = NAN;
  bool have_rmse = false;
  double c2zero = 0.0;
  // hipGraph step blocks: runs of unsampled steps, captured once per length
  bool use_graphs = true;
  std::vector<std::pair<int, hipGraphExec_t>> graphs;
};

// ---------------------------------------------------------------------------
// parameters
// ---------------------------------------------------------------------------
extern "C" int xpfc_params_default(xpfc_params* out) {
  if (!out) return fail(XPFC_ERR_ARG, "null params");
  std::memset(out, 0, sizeof(*out));
  out->dim = 2;
  out->n[0] = out->n[1] = 256;
  out->n[2] = 1;
  out->L[0] = out->L[1] = 32.0;
  out->L[2] = 1.0;
  out->dt = 0.5;
  out->Mo = 1.0;
  out->eps = 0.0;
  out->a = 1.0;
  out->b = 1.0 / 3.0;
  out->sigma = 0.1;
  out->psi0 = 0.05;
  // initialize.cu:10-17
  out->npeaks = 2;
  out->peak_k[0] = 2.0 * M_PI;
  out->peak_k[1] = 2.0 * M_PI * std::sqrt(2.0);
  out->peak_alpha[0] = out->peak_alpha[1] = 1.0;
  out->peak_rho_beta[0] = 1.0 * 4.0;
  out->peak_rho_beta[1] = (1.0 / std::sqrt(2.0)) * 4.0;
  out->dcf_mode = XPFC_DCF_SUM;
  out->precision = XPFC_FP64;
  out->t_sample = 100;
  out->int_Mo_residual = 1;
  out->iconfig = 2;
  return XPFC_OK;
}

namespace {
// fscanf + fgets(tt, 80) pairs exactly as read_input.cu:19-115
bool rd_line_rest(FILE* f) {
  char tt[80];
  return std::fgets(tt, 80, f) != nullptr;
}
}  // namespace

extern "C" int xpfc_read_input(const char* path, xpfc_params* out) {
  if (!path || !out) return fail(XPFC_ERR_ARG, "null argument");
  FILE* f = std::fopen(path, "r");
  if (!f) return fail(XPFC_ERR_IO, "cannot open %s", path);
  xpfc_params p;
  xpfc_params_default(&p);
  double a_in = 0, b_in = 0;
  int Lx = 0, Ly = 0, Mx = 0, My = 0;
  bool ok = true;
  ok = ok && std::fscanf(f, "%d", &p.no_val) == 1 && rd_line_rest(f);
  ok = ok && std::fscanf(f, "%lf %lf", &p.psi0, &p.sigma) == 2 && rd_line_rest(f);
  ok = ok && std::fscanf(f, "%lf %lf %lf %lf", &a_in, &b_in, &p.eps, &p.Mo) == 4 && rd_line_rest(f);
  ok = ok && rd_line_rest(f);
  ok = ok && std::fscanf(f, "%d %d", &Lx, &Ly) == 2 && rd_line_rest(f);
  ok = ok && std::fscanf(f, "%d %d", &Mx, &My) == 2 && rd_line_rest(f);
  ok = ok && rd_line_rest(f);
  ok = ok && std::fscanf(f, "%d", &p.t_final) == 1 && rd_line_rest(f);
  ok = ok && std::fscanf(f, "%lf", &p.dt) == 1 && rd_line_rest(f);
  ok = ok && std::fscanf(f, "%d", &p.t_sample) == 1 && rd_line_rest(f);
  ok = ok && rd_line_rest(f);
  ok = ok && std::fscanf(f, "%d", &p.iconfig) == 1 && rd_line_rest(f);
  ok = ok && std::fscanf(f, "%lf %lf", &p.A_10, &p.A_11) == 2 && rd_line_rest(f);
  ok = ok && std::fscanf(f, "%lf", &p.A_tri) == 1 && rd_line_rest(f);
  ok = ok && rd_line_rest(f);
  ok = ok && std::fscanf(f, "%d", &p.iseed) == 1 && rd_line_rest(f);
  ok = ok && std::fscanf(f, "%d", &p.ishape) == 1 && rd_line_rest(f);
  if (ok) {
    ok = std::fscanf(f, "%lf %lf %lf", &p.h, &p.k, &p.roa) == 3;
    rd_line_rest(f);
  }
  if (!ok) {
    std::fclose(f);
    return fail(XPFC_ERR_PARSE, "%s: malformed pfc.input", path);
  }
  // optional extension lines "key value..." after the 18 reference lines
  char line[512];
  while (std::fgets(line, sizeof(line), f)) {
    char key[64];
    if (std::sscanf(line, "%63s", key) != 1 || key[0] == '#') continue;
    const char* rest = line + std::strlen(key);
    if (!std::strcmp(key, "dim")) std::sscanf(rest, "%d", &p.dim);
    else if (!std::strcmp(key, "Lz")) std::sscanf(rest, "%lf", &p.L[2]);
    else if (!std::strcmp(key, "Mz")) std::sscanf(rest, "%d", &p.n[2]);
    else if (!std::strcmp(key, "precision")) {
      char v[32] = {0};
      std::sscanf(rest, "%31s", v);
      p.precision = (!std::strcmp(v, "fp32") || !std::strcmp(v, "1")) ? XPFC_FP32_MIXED : XPFC_FP64;
    } else if (!std::strcmp(key, "dcf")) {
      char v[32] = {0};
      std::sscanf(rest, "%31s", v);
      p.dcf_mode = !std::strcmp(v, "max") ? XPFC_DCF_MAX : XPFC_DCF_SUM;
    } else if (!std::strcmp(key, "peaks")) {
      // peaks n k1 a1 rb1 k2 a2 rb2 ...
      std::istringstream is(rest);
      int n = 0;
      is >> n;
      if (n < 1 || n > XPFC_MAX_PEAKS) {
        std::fclose(f);
        return fail(XPFC_ERR_PARSE, "peaks: count must be 1..%d", XPFC_MAX_PEAKS);
      }
      p.npeaks = n;
      for (int i = 0; i < n; ++i) is >> p.peak_k[i] >> p.peak_alpha[i] >> p.peak_rho_beta[i];
      if (!is) {
        std::fclose(f);
        return fail(XPFC_ERR_PARSE, "peaks: expected 3 numbers per peak");
      }
    } else {
      std::fclose(f);
      return fail(XPFC_ERR_PARSE, "unknown extension key '%s'", key);
    }
  }
  std::fclose(f);
  if (p.iconfig != 0 && p.iconfig != 1 && p.iconfig != 2 && p.iconfig != 10)
    return fail(XPFC_ERR_PARSE, "Invalid Initital Configuration %d", p.iconfig);
  if (p.iseed != 0 && p.iseed != 1) return fail(XPFC_ERR_PARSE, "Invalid seed flag %d", p.iseed);
  if (p.ishape != 0 && p.ishape != 1) return fail(XPFC_ERR_PARSE, "Invalid Initital Seed Shape %d", p.ishape);
  if (a_in == 0.0 || b_in == 0.0) return fail(XPFC_ERR_PARSE, "a_in and b_in must be non-zero");
  p.a = 1.0 / a_in;  // read_input.cu:28
  p.b = 1.0 / b_in;
  // initialize.cu:23-29: triangular boxes are scaled by a_tri with int truncation
  if (p.iconfig == 0) {
    const double a_tri = 2.0 / std::sqrt(3.0);
    Lx = (int)(Lx * a_tri);
    Ly = (int)(Ly * a_tri);
  }
  p.L[0] = Lx;
  p.L[1] = Ly;
  p.n[0] = Mx;
  p.n[1] = My;
  if (p.dim == 2) {
    p.n[2] = 1;
    p.L[2] = 1.0;
  }
  *out = p;
  return XPFC_OK;
}

extern "C" long long xpfc_reference_steps(const xpfc_params* p) {
  // initialize.cu:60: N_t_steps = t/dt + 1 (double); main.cu:306 loops n < N_t_steps
  const double nt = p->t_final / p->dt + 1.0;
  long long n = 0;
  while ((double)n < nt) ++n;
  return n;
}

// ---------------------------------------------------------------------------
// Philox4x32-10 and randn
// ---------------------------------------------------------------------------
extern "C" void xpfc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

extern "C" int xpfc_randn(unsigned long long seed, unsigned long long first, long long count,
                          double* out) {
  if (count < 0 || (count > 0 && !out)) return fail(XPFC_ERR_ARG, "bad randn arguments");
  const uint32_t key[2] = {(uint32_t)(seed & 0xFFFFFFFFu), (uint32_t)(seed >> 32)};
  const double two_pi = 2.0 * M_PI;
  const double s32 = 1.0 / 4294967296.0;
  for (long long i = 0; i < count; ++i) {
    const unsigned long long c = first + (unsigned long long)i;
    const uint32_t ctr[4] = {(uint32_t)(c & 0xFFFFFFFFu), (uint32_t)(c >> 32), 0u, 0u};
    uint32_t r[4];
    xpfc_philox4x32_10(ctr, key, r);
    const double u0 = ((double)r[0] + 0.5) * s32;
    const double u1 = ((double)r[1] + 0.5) * s32;
    out[i] = std::sqrt(-2.0 * std::log(u0)) * std::cos(two_pi * u1);
  }
  return XPFC_OK;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
namespace {

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

double c2hat_host(double kmag, const xpfc_params& p) {
  double r = 0.0;
  for (int i = 0; i < p.npeaks; ++i) {
    const double amp = std::exp(-(p.sigma * p.peak_k[i]) * (p.sigma * p.peak_k[i]) /
                                (2.0 * p.peak_rho_beta[i]));
    const double d = kmag - p.peak_k[i];
    const double term = amp * std::exp(-(d * d) / (2.0 * p.peak_alpha[i] * p.peak_alpha[i]));
    r = p.dcf_mode == XPFC_DCF_MAX ? std::max(r, term) : r + term;
  }
  return r;
}

// widest valid tile (16 columns FP32 / 8 FP64 keeps every row access 128 bytes)
int default_tc(int prec, int L, const Segs& s) {
  const int pref = prec == 1 ? 16 : 8;
  for (int tc = pref; tc >= 4; tc /= 2)
    if (strided_blocks(prec, L, tc, s) >= 0) return tc;
  return -1;
}

// Captured step blocks bake in buffer pointers, pass variants and the mean
// density (a kernel argument): drop them whenever one of those changes.
void drop_graphs(xpfc_ctx* c) {
  for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.second);
  c->graphs.clear();
}

void free_all(xpfc_ctx* c) {
  drop_graphs(c);
  void* ptrs[] = {c->psi, c->work, c->aux, c->tw0, c->tw1, c->twz, c->twp,
                  c->px, c->pz, c->sums};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  if (c->stream) (void)hipStreamDestroy(c->stream);
}

}  // namespace

extern "C" int xpfc_create(const xpfc_params* pin, xpfc_ctx** out) {
  if (!pin || !out) return fail(XPFC_ERR_ARG, "null argument");
  *out = nullptr;
  xpfc_params p = *pin;
  if (p.dim != 2 && p.dim != 3) return fail(XPFC_ERR_ARG, "dim must be 2 or 3");
  if (p.dim == 2) {
    p.n[2] = 1;
    p.L[2] = 1.0;
  }
  const int nd = p.dim;
  for (int d = 0; d < nd; ++d) {
    if (!is_pow2(p.n[d]) || p.n[d] < 8 || p.n[d] > 1024)
      return fail(XPFC_ERR_ARG, "grid extent n[%d]=%d must be a power of two in [8, 1024]", d, p.n[d]);
    if (!(p.L[d] > 0.0)) return fail(XPFC_ERR_ARG, "box length L[%d] must be positive", d);
  }
  if (p.npeaks < 1 || p.npeaks > XPFC_MAX_PEAKS) return fail(XPFC_ERR_ARG, "npeaks out of range");
  if (p.precision != XPFC_FP64 && p.precision != XPFC_FP32_MIXED)
    return fail(XPFC_ERR_ARG, "bad precision");
  if (!(p.dt > 0.0)) return fail(XPFC_ERR_ARG, "dt must be positive");

  xpfc_ctx* c = new (std::nothrow) xpfc_ctx();
  if (!c) return fail(XPFC_ERR_NOMEM, "host allocation failed");
  c->p = p;
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
  c->rsize = c->prec == 1 ? 4 : 8;

  int rc = XPFC_OK;
  auto bail = [&](int r) {
    free_all(c);
    delete c;
    return r;
  };
  if (hipSetDevice(p.device) != hipSuccess) return bail(fail(XPFC_ERR_DEVICE, "hipSetDevice(%d) failed", p.device));
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(XPFC_ERR_DEVICE, "stream creation failed"));
  const size_t sbytes = (size_t)(c->nmain + c->nny) * c->csize;
  if (hipMalloc(&c->psi, sbytes) != hipSuccess || hipMalloc(&c->work, sbytes) != hipSuccess ||
      hipMalloc(&c->aux, sbytes) != hipSuccess)
    return bail(fail(XPFC_ERR_NOMEM, "device allocation of %zu bytes x3 failed", sbytes));
  if ((rc = upload_twiddles(c, &c->tw0, c->n0, c->n0)) != XPFC_OK) return bail(rc);
  if ((rc = upload_twiddles(c, &c->tw1, c->n1, c->n1)) != XPFC_OK) return bail(rc);
  if ((rc = upload_twiddles(c, &c->twz, c->H, c->H)) != XPFC_OK) return bail(rc);
  if ((rc = upload_twiddles(c, &c->twp, c->H, c->n2)) != XPFC_OK) return bail(rc);

  // strided segments
  Segs sx{};
  sx.s[0] = Seg{0, 0, (int)(c->n1 * c->H), 1, (int)(c->n1 * c->H), 1, 0, 0};
  sx.s[1] = Seg{c->nmain, 0, c->n1, 1, c->n1, 1, 0, 1};
  Segs sy{};
  sy.s[0] = Seg{0, (long long)c->n1 * c->H, c->H, 1, c->H, c->n0, 0, 0};
  sy.s[1] = Seg{c->nmain, 0, 1, c->n1, c->n0, 1, 0, 1};
  c->segX = sx;
  c->segY = sy;
  // default variants: the widest tile that fits, 256-thread real-space blocks
  c->tc_x = default_tc(c->prec, c->n0, sx);
  c->tc_y = c->n1 > 1 ? default_tc(c->prec, c->n1, sy) : 0;
  if (c->tc_x <= 0 || (c->n1 > 1 && c->tc_y <= 0)) return bail(fail(XPFC_ERR_ARG, "unsupported grid"));
  c->ntz = zpass_blocks(c->prec, c->H, 256, 1) > 0 ? 256 : 512;
  // partial buffers sized for the largest variant (narrowest tile, smallest block)
  c->npx_cap = 4096;
  for (int tc : {4, 8, 16}) c->npx_cap = std::max(c->npx_cap, strided_blocks(c->prec, c->n0, tc, sx));
  c->npz_cap = 1;
  for (int nz : {64, 128, 256, 512}) c->npz_cap = std::max(c->npz_cap, zpass_blocks(c->prec, c->H, nz, (long long)c->n0 * c->n1));
  if (hipMalloc((void**)&c->px, (size_t)c->npx_cap * XPART * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&c->pz, (size_t)c->npz_cap * ZPART * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&c->sums, (size_t)SUM_SLOTS * NSUMS * sizeof(double)) != hipSuccess)
    return bail(fail(XPFC_ERR_NOMEM, "partial-sum allocation failed"));

  // geometry and physics
  c->g.n0 = c->n0;
  c->g.n1 = c->n1;
  c->g.H = c->H;
  c->g.dk0 = 2.0 * M_PI / L0;
  c->g.dk1 = nd == 3 ? 2.0 * M_PI / L1 : 0.0;
  c->g.dk2 = 2.0 * M_PI / L2;
  Phys& ph = c->ph;
  ph.dtM = p.dt * p.Mo;
  ph.one_m_eps = 1.0 - p.eps;
  ph.inv_dt = 1.0 / p.dt;
  ph.Mres = p.int_Mo_residual ? (double)(int)p.Mo : p.Mo;
  ph.npk = p.npeaks;
  ph.dcf_max = p.dcf_mode == XPFC_DCF_MAX;
  // a peak term below exp(-ecut) is dropped: 1e-10 (FP32 fields) / 4e-18 (FP64)
  // of the denominator, which is >= 1
  ph.ecut = c->prec == 1 ? 23.0 : 40.0;
  double kc = 0.0;
  for (int i = 0; i < p.npeaks; ++i) {
    ph.pk[i] = p.peak_k[i];
    ph.inv2a2[i] = 1.0 / (2.0 * p.peak_alpha[i] * p.peak_alpha[i]);
    ph.amp[i] = std::exp(-(p.sigma * p.peak_k[i]) * (p.sigma * p.peak_k[i]) / (2.0 * p.peak_rho_beta[i]));
    kc = std::max(kc, p.peak_k[i] + std::sqrt(ph.ecut / ph.inv2a2[i]));
  }
  ph.kcut2 = kc * kc;
  c->c2zero = c2hat_host(0.0, p);
  c->za.a = p.a;
  c->za.b = p.b;
  c->za.scale = 1.0 / (double)c->N;
  c->za.mean = p.psi0;
  c->mean = p.psi0;
  const char* eg = std::getenv("XPFC_GRAPHS");
  c->use_graphs = !(eg && eg[0] == '0');
  *out = c;
  return XPFC_OK;
}

extern "C" void xpfc_destroy(xpfc_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_all(c);
  delete c;
}

extern "C" int xpfc_get_params(const xpfc_ctx* c, xpfc_params* out) {
  if (!c || !out) return fail(XPFC_ERR_ARG, "null argument");
  *out = c->p;
  return XPFC_OK;
}

extern "C" long long xpfc_num_cells(const xpfc_ctx* c) { return c ? c->N : 0; }
extern "C" long long xpfc_current_step(const xpfc_ctx* c) { return c ? c->step : 0; }
extern "C" size_t xpfc_real_bytes(const xpfc_ctx* c) { return c ? (size_t)c->N * c->rsize : 0; }
extern "C" size_t xpfc_spectrum_bytes(const xpfc_ctx* c) {
  return c ? (size_t)c->n0 * c->n1 * (c->H + 1) * c->csize : 0;
}

// ---------------------------------------------------------------------------
// field I/O
// ---------------------------------------------------------------------------
namespace {

void* ny_of(const xpfc_ctx* c, void* spec) {
  return (char*)spec + (size_t)c->nmain * c->csize;
}

// fluctuation (host FP64) -> aux (working precision, real layout)
int upload_fluct(xpfc_ctx* c, const double* host, double mean) {
  const size_t n = (size_t)c->N;
  if (c->prec == 1) {
    std::vector<float> f(n);
    for (size_t i = 0; i < n; ++i) f[i] = (float)(host[i] - mean);
    HIPCHK(hipMemcpyAsync(c->aux, f.data(), n * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  } else {
    std::vector<double> f(n);
    for (size_t i = 0; i < n; ++i) f[i] = host[i] - mean;
    HIPCHK(hipMemcpyAsync(c->aux, f.data(), n * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return XPFC_OK;
}

// full forward transform of a spectrum-layout buffer whose z axis is already done
int forward_xy(xpfc_ctx* c, void* spec, bool do_x) {
  if (c->n1 > 1) LCHK(launch_strided(c->prec, spec, c->segY, c->n1, c->tc_y, -1, c->tw1, c->stream));
  if (do_x)
    LCHK(launch_strided(c->prec, spec, c->segX, c->n0, c->tc_x & ~XOCC_HINT, -1, c->tw0, c->stream));
  return XPFC_OK;
}

int inverse_xy(xpfc_ctx* c, void* spec) {
  LCHK(launch_strided(c->prec, spec, c->segX, c->n0, c->tc_x & ~XOCC_HINT, +1, c->tw0, c->stream));
  if (c->n1 > 1) LCHK(launch_strided(c->prec, spec, c->segY, c->n1, c->tc_y, +1, c->tw1, c->stream));
  return XPFC_OK;
}

// psi state spectrum -> real field (working precision) in aux, optional sums
int psi_to_real(xpfc_ctx* c, bool sums) {
  const size_t sbytes = (size_t)(c->nmain + c->nny) * c->csize;
  HIPCHK(hipMemcpyAsync(c->aux, c->psi, sbytes, hipMemcpyDeviceToDevice, c->stream));
  int rc = inverse_xy(c, c->aux);
  if (rc) return rc;
  LCHK(launch_zpass(c->prec, Z_C2R, c->aux, ny_of(c, c->aux), c->aux, nullptr,
                    (long long)c->n0 * c->n1, c->H, 256, c->za, c->twz, c->twp, c->pz, sums, c->stream));
  return XPFC_OK;
}

}  // namespace

extern "C" int xpfc_set_field(xpfc_ctx* c, const double* host) {
  if (!c || !host) return fail(XPFC_ERR_ARG, "null argument");
  // Avg_calc.cu:3-18 mean, FP64, row-major order
  double s = 0.0;
  for (long long i = 0; i < c->N; ++i) s += host[i];
  const double mean = s / (double)c->N;
  if (!std::isfinite(mean)) return fail(XPFC_ERR_NONFINITE, "field contains NaN/Inf");
  c->mean = mean;
  c->za.mean = mean;
  int rc = upload_fluct(c, host, mean);
  if (rc) return rc;
  const long long nl = (long long)c->n0 * c->n1;
  // psi_k = FFT(psi - mean); the k = 0 slot is then zero by construction
  LCHK(launch_zpass(c->prec, Z_R2C_FLUCT, c->aux, nullptr, c->psi, ny_of(c, c->psi), nl, c->H,
                    256, c->za, c->twz, c->twp, nullptr, false, c->stream));
  if ((rc = forward_xy(c, c->psi, true))) return rc;
  HIPCHK(hipMemsetAsync(c->psi, 0, c->csize, c->stream));
  // work = FFT_yz(N(psi)): the nonlinear term, all axes but axis 0
  LCHK(launch_zpass(c->prec, Z_R2C_NL, c->aux, nullptr, c->work, ny_of(c, c->work), nl, c->H,
                    256, c->za, c->twz, c->twp, nullptr, false, c->stream));
  if ((rc = forward_xy(c, c->work, false))) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  c->have_field = true;
  drop_graphs(c);  // the mean density is a captured kernel argument
  return XPFC_OK;
}

extern "C" int xpfc_get_field(xpfc_ctx* c, double* host) {
  if (!c || !host) return fail(XPFC_ERR_ARG, "null argument");
  if (!c->have_field) return fail(XPFC_ERR_STATE, "no field set");
  int rc = psi_to_real(c, false);
  if (rc) return rc;
  const size_t n = (size_t)c->N;
  if (c->prec == 1) {
    std::vector<float> f(n);
    HIPCHK(hipMemcpyAsync(f.data(), c->aux, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < n; ++i) host[i] = (double)f[i];
  } else {
    HIPCHK(hipMemcpyAsync(host, c->aux, n * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return XPFC_OK;
}

// ---------------------------------------------------------------------------
// initial conditions
// ---------------------------------------------------------------------------
namespace {

std::vector<double> linspace_ref(double x1, double x2, int N) {
  // linspace.cu:5-18 (without the one-past-the-end write)
  std::vector<double> v(N);
  const double du = x2 / (N - 1);
  v[0] = x1;
  for (int i = 0; i + 1 < N; ++i) v[i + 1] = v[i] + du;
  return v;
}

void fix_mean(std::vector<double>& psi, double psi0) {
  // Avg_calc.cu:21-43
  double s = 0.0;
  for (double x : psi) s += x;
  const double m = s / (double)psi.size();
  double diff = 0.0;
  if (m != psi0) diff = psi0 - m;
  for (double& x : psi) x += diff;
}

int read_vtk_file(const char* path, int* nx, int* ny, double* spacing, std::vector<double>* vals);

}  // namespace

extern "C" int xpfc_initialize(xpfc_ctx* c) {
  if (!c) return fail(XPFC_ERR_ARG, "null context");
  const xpfc_params& p = c->p;
  std::vector<double> psi((size_t)c->N, p.psi0);
  if (p.iconfig == 10) {
    int nx = 0, ny = 0;
    double sp[3];
    std::vector<double> v;
    int rc = read_vtk_file("initial_density_field.vtk", &nx, &ny, sp, &v);
    if (rc) return rc;
    if ((long long)v.size() != c->N) return fail(XPFC_ERR_ARG, "VTK field size does not match the grid");
    psi = v;  // read_input.cu:208-218 fills psi[i*My + j] in file order
  } else if (p.dim == 2) {
    const int Mx = p.n[0], My = p.n[1];
    const auto x = linspace_ref(0.0, p.L[0], Mx);
    const auto y = linspace_ref(0.0, p.L[1], My);
    const double a_tri = 2.0 / std::sqrt(3.0), a_sq = 1.0;
    const double k_sq = 2.0 * M_PI / a_sq, k_tri = 2.0 * M_PI / a_tri;
    double r = 0.0;
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
    if (p.iconfig == 1) r = p.roa * a_sq;
    if (p.iconfig == 0) r = p.roa * a_tri;
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
  return xpfc_set_field(c, psi.data());
}

extern "C" int xpfc_initialize_noise(xpfc_ctx* c, double amplitude, unsigned long long seed) {
  if (!c) return fail(XPFC_ERR_ARG, "null context");
  std::vector<double> psi((size_t)c->N);
  int rc = xpfc_randn(seed, 0, c->N, psi.data());
  if (rc) return rc;
  for (double& v : psi) v = c->p.psi0 + amplitude * v;
  fix_mean(psi, c->p.psi0);
  return xpfc_set_field(c, psi.data());
}

// ---------------------------------------------------------------------------
// the step driver
// ---------------------------------------------------------------------------
namespace {

bool sample_step(const xpfc_ctx* c, long long n) {
  const long long ts = c->p.t_sample;
  return ts > 0 && (n % ts == 0 || (n + 1) % ts == 0);
}

// The four launches of one step (two in 2-D).  With `ev`, an event is recorded
// after each pass (ev[0..3]) for per-pass timing.
int launch_one_step(xpfc_ctx* c, bool sample, double* sums_out, hipEvent_t* ev = nullptr) {
  const long long nl = (long long)c->n0 * c->n1;
  const int nbx = launch_xstep(c->prec, c->work, c->psi, c->segX, c->n0, c->tc_x, c->tw0, c->g,
                               c->ph, c->px, sample, c->stream);
  LCHK(nbx);
  if (ev) HIPCHK(hipEventRecord(ev[0], c->stream));
  if (c->n1 > 1) LCHK(launch_strided(c->prec, c->work, c->segY, c->n1, c->tc_y, +1, c->tw1, c->stream));
  if (ev) HIPCHK(hipEventRecord(ev[1], c->stream));
  LCHK(launch_zpass(c->prec, Z_STEP, c->work, ny_of(c, c->work), c->work, ny_of(c, c->work), nl,
                    c->H, c->ntz, c->za, c->twz, c->twp, c->pz, sample, c->stream));
  if (ev) HIPCHK(hipEventRecord(ev[2], c->stream));
  if (c->n1 > 1) LCHK(launch_strided(c->prec, c->work, c->segY, c->n1, c->tc_y, -1, c->tw1, c->stream));
  if (ev) HIPCHK(hipEventRecord(ev[3], c->stream));
  if (sample)
    LCHK(launch_finalize(c->px, nbx, c->pz, zpass_blocks(c->prec, c->H, c->ntz, nl), sums_out, c->stream));
  return XPFC_OK;
}

// Unsampled steps run as replays of captured hipGraphs of up to GRAPH_BLOCK
// steps (one graph per run length, cached per context).  The captured work is
// the same launch sequence as launch_one_step, so a replay and an eager run
// give identical fields.
constexpr int GRAPH_BLOCK = 50;

int plain_graph(xpfc_ctx* c, int len, hipGraphExec_t* out) {
  for (auto& g : c->graphs)
    if (g.first == len) {
      *out = g.second;
      return XPFC_OK;
    }
  if (c->graphs.size() >= 8) drop_graphs(c);
  hipGraph_t graph = nullptr;
  HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rc = XPFC_OK;
  for (int i = 0; i < len && rc == XPFC_OK; ++i) rc = launch_one_step(c, false, nullptr);
  const hipError_t ec = hipStreamEndCapture(c->stream, &graph);
  if (rc != XPFC_OK || ec != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc != XPFC_OK ? rc : fail(XPFC_ERR_DEVICE, "step capture failed: %s", hipGetErrorString(ec));
  }
  hipGraphExec_t ex = nullptr;
  const hipError_t ei = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) return fail(XPFC_ERR_DEVICE, "graph instantiation failed: %s", hipGetErrorString(ei));
  c->graphs.emplace_back(len, ex);
  *out = ex;
  return XPFC_OK;
}

int run_plain_steps(xpfc_ctx* c, long long run) {
  while (run > 0) {
    if (!c->use_graphs) {
      int rc = launch_one_step(c, false, nullptr);
      if (rc) return rc;
      c->step++;
      --run;
      continue;
    }
    const int len = (int)std::min<long long>(run, GRAPH_BLOCK);
    hipGraphExec_t ex = nullptr;
    int rc = plain_graph(c, len, &ex);
    if (rc) return rc;
    HIPCHK(hipGraphLaunch(ex, c->stream));
    c->step += len;
    run -= len;
  }
  return XPFC_OK;
}

void add_fe_record(xpfc_ctx* c, long long step, double conv, double s1, double s2, double s3, double s4) {
  const xpfc_params& p = c->p;
  const double N = (double)c->N;
  // Parseval: sum_x psi (C2*psi) = (1/N) sum_k c2 |psi_k|^2, k = 0 carried in FP64
  const double conv_x = (conv + c->c2zero * (N * c->mean) * (N * c->mean)) / N;
  const double sf = 0.5 * s2 - (p.a / 3.0) * s3 + (p.b / 4.0) * s4 - 0.5 * conv_x;  // FE_calc.cu:64
  double ff = 1.0;  // initialize.cu:63
  for (int d = 0; d < p.dim; ++d) ff *= p.L[d] / (double)(p.n[d] - 1);
  xpfc_sample r{};
  r.step = step;
  r.has_fe = 1;
  r.F_ref = ff * sf;
  r.f_mean = sf / N;
  r.mean = s1 / N;
  r.rmse = NAN;
  if (!c->samples.empty() && c->samples.back().step == step && !c->samples.back().has_fe) {
    xpfc_sample& b = c->samples.back();
    b.has_fe = 1;
    b.F_ref = r.F_ref;
    b.f_mean = r.f_mean;
    b.mean = r.mean;
  } else {
    c->samples.push_back(r);
  }
}

void add_sse_record(xpfc_ctx* c, long long step, double sse) {
  const double N = (double)c->N;
  const double rmse = std::sqrt(sse) / N;  // sqrt(sum_x R^2 / N), Parseval
  c->last_rmse = rmse;
  c->have_rmse = true;
  if (!c->samples.empty() && c->samples.back().step == step) {
    c->samples.back().has_sse = 1;
    c->samples.back().rmse = rmse;
    return;
  }
  xpfc_sample r{};
  r.step = step;
  r.has_sse = 1;
  r.rmse = rmse;
  r.F_ref = r.f_mean = r.mean = NAN;
  c->samples.push_back(r);
}

int drain(xpfc_ctx* c) {
  if (c->pending.empty()) return XPFC_OK;
  std::vector<double> h((size_t)SUM_SLOTS * NSUMS);
  const int nslot = (int)c->pending.size();
  HIPCHK(hipMemcpyAsync(h.data(), c->sums, (size_t)nslot * NSUMS * sizeof(double),
                        hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  const long long ts = c->p.t_sample;
  bool bad = false;
  for (auto& pr : c->pending) {
    const double* s = &h[(size_t)pr.first * NSUMS];
    const long long n = pr.second;
    for (int k = 0; k < NSUMS; ++k) bad |= !std::isfinite(s[k]);
    if (n % ts == 0) add_sse_record(c, n, s[1]);
    if ((n + 1) % ts == 0) add_fe_record(c, n + 1, s[0], s[2], s[3], s[4], s[5]);
  }
  c->pending.clear();
  if (bad) return fail(XPFC_ERR_NONFINITE, "non-finite value in a sampled reduction");
  return XPFC_OK;
}

int fe_now(xpfc_ctx* c, double* F, double* fm, double* mean_rs);

}  // namespace

// ---------------------------------------------------------------------------
// autotuner: times every variant of each pass on scratch buffers and keeps the
// fastest; results are cached per (precision, grid) for the process.  The
// variants only regroup lines into workgroups, so the choice never changes
// the computed field.
// ---------------------------------------------------------------------------
namespace {
std::mutex g_tune_mu;
std::map<std::array<int, 4>, std::array<int, 3>> g_tune_cache;

int time_pass(xpfc_ctx* c, int which, int v, void* a, void* b, int reps, float* ms) {
  const long long nl = (long long)c->n0 * c->n1;
  auto launch = [&]() -> int {
    if (which == 0)
      return launch_xstep(c->prec, a, b, c->segX, c->n0, v, c->tw0, c->g, c->ph, c->px, false, c->stream);
    if (which == 1) return launch_strided(c->prec, a, c->segY, c->n1, v, -1, c->tw1, c->stream);
    return launch_zpass(c->prec, Z_STEP, a, ny_of(c, a), a, ny_of(c, a), nl, c->H, v, c->za, c->twz,
                        c->twp, c->pz, false, c->stream);
  };
  int r = launch();  // warm-up; also tells whether the variant exists
  if (r == -3) return -3;
  if (r < 0) return fail(XPFC_ERR_DEVICE, "autotune launch failed (%d)", r);
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, c->stream));
  for (int i = 0; i < reps; ++i)
    if (launch() < 0) return fail(XPFC_ERR_DEVICE, "autotune launch failed");
  HIPCHK(hipEventRecord(e1, c->stream));
  HIPCHK(hipEventSynchronize(e1));
  HIPCHK(hipEventElapsedTime(ms, e0, e1));
  *ms /= (float)reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return 0;
}

int autotune(xpfc_ctx* c) {
  const std::array<int, 4> key{c->prec, c->n0, c->n1, c->n2};
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      c->tc_x = it->second[0];
      c->tc_y = it->second[1];
      c->ntz = it->second[2];
      c->tuned = true;
      return XPFC_OK;
    }
  }
  const size_t sbytes = (size_t)(c->nmain + c->nny) * c->csize;
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, sbytes) != hipSuccess || hipMalloc(&b, sbytes) != hipSuccess) {
    if (a) (void)hipFree(a);
    return fail(XPFC_ERR_NOMEM, "autotune scratch allocation failed");
  }
  int rc = XPFC_OK;
  HIPCHK(hipMemsetAsync(a, 0, sbytes, c->stream));
  HIPCHK(hipMemsetAsync(b, 0, sbytes, c->stream));
  const int reps = c->N >= (1LL << 24) ? 4 : 16;
  int best[3] = {c->tc_x, c->tc_y, c->ntz};
  // axis-0: tile width (+XOCC_HINT); axis-1: tile width; real space: block (+ZFREE_REGS)
  const std::vector<int> xs{4, 8, 16, 4 | XOCC_HINT, 8 | XOCC_HINT, 16 | XOCC_HINT};
  const std::vector<int> ys{4, 8, 16};
  const std::vector<int> zs{64, 128, 256, 512, 64 | ZFREE_REGS, 128 | ZFREE_REGS,
                            256 | ZFREE_REGS, 512 | ZFREE_REGS};
  for (int which = 0; which < 3 && rc == XPFC_OK; ++which) {
    if (which == 1 && c->n1 <= 1) continue;
    float bt = 1e30f;
    for (int v : (which == 0 ? xs : which == 1 ? ys : zs)) {
      float ms = 0.f;
      const int r = time_pass(c, which, v, a, b, reps, &ms);
      if (r == -3) continue;
      if (r) { rc = r; break; }
      if (ms < bt) { bt = ms; best[which] = v; }
    }
  }
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(a);
  (void)hipFree(b);
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
  const long long nl = (long long)c->n0 * c->n1;
  if (v[0] > 0 && strided_blocks(c->prec, c->n0, v[0], c->segX) < 0)
    return fail(XPFC_ERR_ARG, "tile width %d not available for n0 = %d", v[0], c->n0);
  if (c->n1 > 1 && v[1] > 0 && strided_blocks(c->prec, c->n1, v[1], c->segY) < 0)
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

extern "C" int xpfc_set_graphs(xpfc_ctx* c, int on) {
  if (!c) return fail(XPFC_ERR_ARG, "null context");
  c->use_graphs = on != 0;
  if (!c->use_graphs) drop_graphs(c);
  return XPFC_OK;
}

extern "C" int xpfc_get_variants(const xpfc_ctx* c, int* v) {
  if (!c || !v) return fail(XPFC_ERR_ARG, "null argument");
  v[0] = c->tc_x;
  v[1] = c->tc_y;
  v[2] = c->ntz;
  return XPFC_OK;
}

extern "C" int xpfc_step(xpfc_ctx* c, int nsteps) {
  if (!c) return fail(XPFC_ERR_ARG, "null context");
  if (!c->have_field) return fail(XPFC_ERR_STATE, "no field set");
  if (nsteps < 0) return fail(XPFC_ERR_ARG, "nsteps < 0");
  if (!c->tuned && autotune_enabled()) {
    int rc = autotune(c);
    if (rc) return rc;
  }
  if (c->step == 0 && c->p.t_sample > 0 && c->samples.empty()) {
    // main.cu:286: FE of the initial field
    double F, fm, mrs;
    int rc = fe_now(c, &F, &fm, &mrs);
    if (rc) return rc;
    xpfc_sample r{};
    r.step = 0;
    r.has_fe = 1;
    r.F_ref = F;
    r.f_mean = fm;
    r.mean = mrs;
    r.rmse = NAN;
    c->samples.push_back(r);
  }
  long long left = nsteps;
  while (left > 0) {
    const long long n = c->step;
    if (sample_step(c, n)) {
      if ((int)c->pending.size() == SUM_SLOTS) {
        int rc = drain(c);
        if (rc) return rc;
      }
      const int s = (int)c->pending.size();
      double* slot = c->sums + (size_t)s * NSUMS;
      c->pending.emplace_back(s, n);
      int rc = launch_one_step(c, true, slot);
      if (rc) return rc;
      c->step++;
      --left;
      continue;
    }
    // a run of unsampled steps up to the next sampled one
    long long run = 1;
    while (run < left && !sample_step(c, n + run)) ++run;
    int rc = run_plain_steps(c, run);
    if (rc) return rc;
    left -= run;
  }
  return drain(c);
}

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
  return rc;
}

extern "C" int xpfc_time_passes(xpfc_ctx* c, int nsteps, double* ms) {
  if (!c || !ms || nsteps <= 0) return fail(XPFC_ERR_ARG, "bad arguments");
  if (!c->have_field) return fail(XPFC_ERR_STATE, "no field set");
  if (!c->tuned && autotune_enabled()) {
    int rc = autotune(c);
    if (rc) return rc;
  }
  std::vector<hipEvent_t> ev((size_t)nsteps * 4 + 1);
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  HIPCHK(hipEventRecord(ev[0], c->stream));
  for (int i = 0; i < nsteps; ++i) {
    int rc = launch_one_step(c, false, nullptr, &ev[(size_t)i * 4 + 1]);
    if (rc) return rc;
    c->step++;
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

extern "C" int xpfc_sync(xpfc_ctx* c) {
  if (!c) return fail(XPFC_ERR_ARG, "null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  return XPFC_OK;
}

extern "C" int xpfc_num_samples(const xpfc_ctx* c) { return c ? (int)c->samples.size() : 0; }
extern "C" int xpfc_get_samples(const xpfc_ctx* c, xpfc_sample* out, int max) {
  if (!c || (max > 0 && !out)) return fail(XPFC_ERR_ARG, "null argument");
  const int n = std::min(max, (int)c->samples.size());
  for (int i = 0; i < n; ++i) out[i] = c->samples[i];
  return n;
}
extern "C" void xpfc_clear_samples(xpfc_ctx* c) {
  if (c) c->samples.clear();
}

// ---------------------------------------------------------------------------
// diagnostics
// ---------------------------------------------------------------------------
namespace {
int fe_now(xpfc_ctx* c, double* F, double* fm, double* mean_rs) {
  int nb = 0;
  LCHK(launch_kspace_sums(c->prec, c->psi, c->nmain, c->nny, c->g, c->ph, c->px, &nb, c->stream));
  int rc = psi_to_real(c, true);
  if (rc) return rc;
  double* slot = c->sums + (size_t)(SUM_SLOTS - 1) * NSUMS;
  LCHK(launch_finalize(c->px, nb, c->pz, zpass_blocks(c->prec, c->H, 256, (long long)c->n0 * c->n1), slot,
                       c->stream));
  double h[NSUMS];
  HIPCHK(hipMemcpyAsync(h, slot, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (double v : h)
    if (!std::isfinite(v)) return fail(XPFC_ERR_NONFINITE, "non-finite free-energy sum");
  const xpfc_params& p = c->p;
  const double N = (double)c->N;
  const double conv_x = (h[0] + c->c2zero * (N * c->mean) * (N * c->mean)) / N;
  const double sf = 0.5 * h[3] - (p.a / 3.0) * h[4] + (p.b / 4.0) * h[5] - 0.5 * conv_x;
  double ff = 1.0;
  for (int d = 0; d < p.dim; ++d) ff *= p.L[d] / (double)(p.n[d] - 1);
  *F = ff * sf;
  *fm = sf / N;
  *mean_rs = h[2] / N;
  return XPFC_OK;
}
}  // namespace

extern "C" int xpfc_fe_calc(xpfc_ctx* c, double* F_ref, double* f_mean) {
  if (!c || !F_ref || !f_mean) return fail(XPFC_ERR_ARG, "null argument");
  if (!c->have_field) return fail(XPFC_ERR_STATE, "no field set");
  double m;
  return fe_now(c, F_ref, f_mean, &m);
}

extern "C" int xpfc_avg_calc(xpfc_ctx* c, double* mean) {
  if (!c || !mean) return fail(XPFC_ERR_ARG, "null argument");
  if (!c->have_field) return fail(XPFC_ERR_STATE, "no field set");
  *mean = c->mean;
  return XPFC_OK;
}

extern "C" int xpfc_sse_calc(xpfc_ctx* c, double* rmse) {
  if (!c || !rmse) return fail(XPFC_ERR_ARG, "null argument");
  if (!c->have_rmse) return fail(XPFC_ERR_STATE, "no sampled step yet");
  *rmse = c->last_rmse;
  return XPFC_OK;
}

// ---------------------------------------------------------------------------
// operator level
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

#define OPCHK(c)                                                      \
  do {                                                                \
    if (!(c)) return fail(XPFC_ERR_ARG, "null context");              \
  } while (0)

extern "C" int xpfc_power_functions(xpfc_ctx* c, int p, const void* src, void* dst) {
  OPCHK(c);
  if (!src || !dst || p < 0) return fail(XPFC_ERR_ARG, "bad power_functions arguments");
  LCHK(launch_power(c->prec, src, dst, c->N, p, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return XPFC_OK;
}

extern "C" int xpfc_convolve_fields(xpfc_ctx* c, const void* psi_k, void* out) {
  OPCHK(c);
  if (!psi_k || !out) return fail(XPFC_ERR_ARG, "null buffer");
  LCHK(launch_convolve(c->prec, psi_k, out, c->g, c->ph, c->n2, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return XPFC_OK;
}

extern "C" int xpfc_euler_step(xpfc_ctx* c, const void* psi_k, const void* p2_k, const void* p3_k,
                               void* out) {
  OPCHK(c);
  if (!psi_k || !p2_k || !p3_k || !out) return fail(XPFC_ERR_ARG, "null buffer");
  LCHK(launch_euler(c->prec, psi_k, p2_k, p3_k, out, c->g, c->ph, c->p.a, c->p.b, c->n2, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return XPFC_OK;
}

extern "C" int xpfc_residual_calc(xpfc_ctx* c, const void* kp1, const void* k0, const void* p2,
                                  const void* p3, void* out) {
  OPCHK(c);
  if (!kp1 || !k0 || !p2 || !p3 || !out) return fail(XPFC_ERR_ARG, "null buffer");
  LCHK(launch_residual(c->prec, kp1, k0, p2, p3, out, c->g, c->ph, c->p.a, c->p.b, c->n2, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return XPFC_OK;
}

extern "C" int xpfc_scaling_step(xpfc_ctx* c, void* buf, int spectral) {
  OPCHK(c);
  if (!buf) return fail(XPFC_ERR_ARG, "null buffer");
  const long long n = spectral ? 2LL * c->n0 * c->n1 * (c->H + 1) : c->N;
  LCHK(launch_scale(c->prec, buf, n, 1.0 / (double)c->N, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return XPFC_OK;
}

extern "C" int xpfc_fft_r2c(xpfc_ctx* c, const void* real_in, void* spec_out) {
  OPCHK(c);
  if (!real_in || !spec_out) return fail(XPFC_ERR_ARG, "null buffer");
  const long long nl = (long long)c->n0 * c->n1;
  ZArgs za = c->za;
  za.mean = 0.0;
  LCHK(launch_zpass(c->prec, Z_R2C_FLUCT, real_in, nullptr, c->aux, ny_of(c, c->aux), nl, c->H, 256, za,
                    c->twz, c->twp, nullptr, false, c->stream));
  int rc = forward_xy(c, c->aux, true);
  if (rc) return rc;
  LCHK(launch_unpack(c->prec, c->aux, spec_out, (int)nl, c->H, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return XPFC_OK;
}

extern "C" int xpfc_fft_c2r(xpfc_ctx* c, const void* spec_in, void* real_out) {
  OPCHK(c);
  if (!spec_in || !real_out) return fail(XPFC_ERR_ARG, "null buffer");
  const long long nl = (long long)c->n0 * c->n1;
  LCHK(launch_pack(c->prec, spec_in, c->aux, (int)nl, c->H, c->stream));
  int rc = inverse_xy(c, c->aux);
  if (rc) return rc;
  ZArgs za = c->za;
  za.mean = 0.0;
  za.scale = 1.0;  // unnormalised, as cuFFT
  LCHK(launch_zpass(c->prec, Z_C2R, c->aux, ny_of(c, c->aux), real_out, nullptr, nl, c->H, 256, za,
                    c->twz, c->twp, nullptr, false, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return XPFC_OK;
}

// ---------------------------------------------------------------------------
// VTK
// ---------------------------------------------------------------------------
extern "C" int xpfc_write_vtk(xpfc_ctx* c, const char* path) {
  if (!c || !path) return fail(XPFC_ERR_ARG, "null argument");
  std::vector<double> psi((size_t)c->N);
  int rc = xpfc_get_field(c, psi.data());
  if (rc) return rc;
  std::ofstream f(path);
  if (!f) return fail(XPFC_ERR_IO, "cannot open %s for writing", path);
  const xpfc_params& p = c->p;
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

namespace {
int read_vtk_file(const char* path, int* nx, int* ny, double* spacing, std::vector<double>* vals) {
  FILE* f = std::fopen(path, "r");
  if (!f) return fail(XPFC_ERR_IO, "cannot open %s", path);
  char tt[80];
  int mx = 0, my = 0, mz = 0;
  bool ok = true;
  for (int i = 0; i < 4 && ok; ++i) ok = std::fgets(tt, 80, f) != nullptr;  // header lines
  ok = ok && std::fscanf(f, "DIMENSIONS %d %d %d", &mx, &my, &mz) == 3;
  ok = ok && std::fgets(tt, 80, f) && std::fgets(tt, 80, f);  // rest of DIMENSIONS, ORIGIN
  double dx = 0, dy = 0, dz = 0;
  ok = ok && std::fscanf(f, "SPACING %lf %lf %lf", &dx, &dy, &dz) == 3;
  for (int i = 0; i < 4 && ok; ++i) ok = std::fgets(tt, 80, f) != nullptr;  // rest, POINT_DATA, SCALARS, LOOKUP
  if (!ok || mx <= 0 || my <= 0 || mz != 1) {
    std::fclose(f);
    return fail(XPFC_ERR_PARSE, "%s: not a 2-D legacy VTK structured-points file", path);
  }
  vals->resize((size_t)mx * my);
  for (size_t i = 0; i < vals->size(); ++i) {
    if (!std::fgets(tt, sizeof(tt), f) || std::sscanf(tt, "%lf", &(*vals)[i]) != 1) {
      std::fclose(f);
      return fail(XPFC_ERR_PARSE, "%s: unexpected end of data at index %zu", path, i);
    }
  }
  std::fclose(f);
  *nx = mx;
  *ny = my;
  if (spacing) {
    spacing[0] = dx;
    spacing[1] = dy;
    spacing[2] = dz;
  }
  return XPFC_OK;
}
}  // namespace

extern "C" int xpfc_read_vtk(const char* path, int* nx, int* ny, double* spacing, double* host_real,
                             long long capacity) {
  if (!path || !nx || !ny) return fail(XPFC_ERR_ARG, "null argument");
  std::vector<double> v;
  int rc = read_vtk_file(path, nx, ny, spacing, &v);
  if (rc) return rc;
  if (host_real) {
    if ((long long)v.size() > capacity) return fail(XPFC_ERR_ARG, "capacity too small");
    std::copy(v.begin(), v.end(), host_real);
  }
  return XPFC_OK;
}

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
extern "C" const char* xpfc_strerror(int s) {
  switch (s) {
    case XPFC_OK: return "ok";
    case XPFC_ERR_ARG: return "invalid argument";
    case XPFC_ERR_IO: return "I/O error";
    case XPFC_ERR_PARSE: return "parse error";
    case XPFC_ERR_DEVICE: return "device error";
    case XPFC_ERR_NOMEM: return "out of memory";
    case XPFC_ERR_STATE: return "invalid state";
    case XPFC_ERR_NONFINITE: return "non-finite result";
    default: return "unknown status";
  }
}

extern "C" const char* xpfc_last_error(void) { return g_last_error.c_str(); }
