// Host-side parts of libxpfc that need no device: parameters and the
// pfc.input reader, the Philox/Box-Muller normal generator, reference
// initial-condition helpers, the VTK reader and error reporting.
#include "xpfc_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

namespace xpfc {

namespace {
thread_local std::string g_last_error;
}

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  // a failed runtime call (for example an allocation refused before a
  // context is torn down) leaves its error pending; clear it so the next
  // launch check does not report it
  (void)hipGetLastError();
  return code;
}

const char* last_error_text() { return g_last_error.c_str(); }

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

int read_vtk_file(const char* path, int* dims, double* spacing, std::vector<double>* vals) {
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
  if (!ok || mx <= 0 || my <= 0 || mz <= 0 || (long long)mx * my * mz > (1LL << 31)) {
    std::fclose(f);
    return fail(XPFC_ERR_PARSE, "%s: not a legacy VTK structured-points file", path);
  }
  vals->resize((size_t)mx * my * mz);
  for (size_t i = 0; i < vals->size(); ++i) {
    if (!std::fgets(tt, sizeof(tt), f) || std::sscanf(tt, "%lf", &(*vals)[i]) != 1) {
      std::fclose(f);
      return fail(XPFC_ERR_PARSE, "%s: unexpected end of data at index %zu", path, i);
    }
  }
  std::fclose(f);
  dims[0] = mx;
  dims[1] = my;
  dims[2] = mz;
  if (spacing) {
    spacing[0] = dx;
    spacing[1] = dy;
    spacing[2] = dz;
  }
  return XPFC_OK;
}
}  // namespace xpfc

using xpfc::fail;

// ---------------------------------------------------------------------------
// parameters
// ---------------------------------------------------------------------------
extern "C" size_t xpfc_params_size(void) { return sizeof(xpfc_params); }

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
    else if (!std::strcmp(key, "restart_order")) std::sscanf(rest, "%d", &p.restart_order);
    else if (!std::strcmp(key, "powf_fp32")) std::sscanf(rest, "%d", &p.powf_fp32);
    else if (!std::strcmp(key, "int_Mo_residual")) std::sscanf(rest, "%d", &p.int_Mo_residual);
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
  if (p.iconfig == 10) {
    // obtain_box_size (read_input.cu:124-157): L = round(M_t * spacing) per
    // axis from the restart file in the working directory
    int dims[3];
    double sp[3];
    std::vector<double> v;
    int rc = xpfc::read_vtk_file("initial_density_field.vtk", dims, sp, &v);
    if (rc) return rc;
    for (int d = 0; d < 3; ++d) p.L[d] = (double)(int)std::lround(dims[d] * sp[d]);
  }
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

namespace {
void randn_range(const uint32_t key[2], unsigned long long first, long long i0, long long i1,
                 double* out) {
  const double two_pi = 2.0 * M_PI;
  const double s32 = 1.0 / 4294967296.0;
  for (long long i = i0; i < i1; ++i) {
    const unsigned long long c = first + (unsigned long long)i;
    const uint32_t ctr[4] = {(uint32_t)(c & 0xFFFFFFFFu), (uint32_t)(c >> 32), 0u, 0u};
    uint32_t r[4];
    xpfc_philox4x32_10(ctr, key, r);
    const double u0 = ((double)r[0] + 0.5) * s32;
    const double u1 = ((double)r[1] + 0.5) * s32;
    out[i] = std::sqrt(-2.0 * std::log(u0)) * std::cos(two_pi * u1);
  }
}
}  // namespace

// Each draw depends only on (seed, cell), so large requests are split over
// host threads with identical results.
extern "C" int xpfc_randn(unsigned long long seed, unsigned long long first, long long count,
                          double* out) {
  if (count < 0 || (count > 0 && !out)) return fail(XPFC_ERR_ARG, "bad randn arguments");
  const uint32_t key[2] = {(uint32_t)(seed & 0xFFFFFFFFu), (uint32_t)(seed >> 32)};
  const long long chunk = 1LL << 19;
  const int nt = (int)std::min<long long>({(count + chunk - 1) / chunk, 16LL,
                                           (long long)std::max(1u, std::thread::hardware_concurrency())});
  if (nt <= 1) {
    randn_range(key, first, 0, count, out);
    return XPFC_OK;
  }
  std::vector<std::thread> th;
  const long long per = (count + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    const long long i0 = t * per, i1 = std::min(count, i0 + per);
    if (i0 < i1) th.emplace_back(randn_range, key, first, i0, i1, out);
  }
  for (auto& x : th) x.join();
  return XPFC_OK;
}

// ---------------------------------------------------------------------------
// VTK restart input (read_input.cu:161-241)
// ---------------------------------------------------------------------------
extern "C" int xpfc_read_vtk(const char* path, int* dims, double* spacing, double* host_real,
                             long long capacity) {
  if (!path || !dims) return fail(XPFC_ERR_ARG, "null argument");
  std::vector<double> v;
  int rc = xpfc::read_vtk_file(path, dims, spacing, &v);
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

extern "C" const char* xpfc_last_error(void) { return xpfc::last_error_text(); }
