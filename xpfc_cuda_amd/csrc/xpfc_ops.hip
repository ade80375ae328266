// Operator-level HIP kernels on natural layouts (real [N], spectrum
// [n0][n1][n2/2+1]): the reference's per-operator kernels, for drop-in
// parity of each operator.  The time step itself does not use them.
#include "xpfc_device.hpp"

#include <algorithm>
#include <mutex>

namespace xpfc {

// ---------------------------------------------------------------------------
// operator-level kernels on natural layouts
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_power(const T* __restrict__ src, T* __restrict__ dst, long long n, int p, int f32) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const T x = src[i];
    if (f32) {  // the reference's FP32 powf (power_functions.cu:47)
      const float xf = (float)x;
      float r = 1.0f;
      for (int q = 0; q < p; ++q) r *= xf;
      dst[i] = (T)r;
    } else {
      T r = (T)1;
      for (int q = 0; q < p; ++q) r *= x;
      dst[i] = r;
    }
  }
}

template <typename T>
__global__ void k_scale(T* __restrict__ buf, long long n, T f) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    buf[i] *= f;
}

// natural spectrum index e -> k^2 and c2
__device__ __forceinline__ void natural_k(long long e, int n2h1, const KGeom& g, const Phys& ph,
                                          double& k2, double& c2) {
  const int kz = (int)(e % n2h1);
  const long long r = e / n2h1;
  const int kb = (int)(r % g.n1);
  const int ka = (int)(r / g.n1);
  const double kx = g.dk0 * signed_index(ka, g.n0), ky = g.dk1 * signed_index(kb, g.n1),
               kzv = g.dk2 * kz;
  k2 = kx * kx + ky * ky + kzv * kzv;
  c2 = (k2 < ph.kcut2) ? c2hat(sqrt(k2), ph) : 0.0;
}

template <typename T>
__global__ void k_convolve(const cpx<T>* __restrict__ psi_k, cpx<T>* __restrict__ out, long long n,
                           int n2h1, KGeom g, Phys ph) {
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    double k2, c2;
    natural_k(e, n2h1, g, ph, k2, c2);
    const cpx<T> v = psi_k[e];
    out[e] = mk<cpx<T>>((T)(c2 * v.x), (T)(c2 * v.y));  // convolve_fields.cu:41-42
  }
}

template <typename T>
__global__ void k_euler(const cpx<T>* __restrict__ psi_k, const cpx<T>* __restrict__ p2,
                        const cpx<T>* __restrict__ p3, cpx<T>* __restrict__ out, long long n,
                        int n2h1, KGeom g, Phys ph, double a, double b) {
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    double k2, c2;
    natural_k(e, n2h1, g, ph, k2, c2);
    const cpx<T> s = psi_k[e], q2 = p2[e], q3 = p3[e];
    // euler_step.cu:50-59
    const double nr = s.x + a * k2 * ph.dtM * q2.x - b * k2 * ph.dtM * q3.x;
    const double ni = s.y + a * k2 * ph.dtM * q2.y - b * k2 * ph.dtM * q3.y;
    const double den = 1.0 + k2 * ph.dtM * ph.one_m_eps - k2 * ph.dtM * c2;
    out[e] = mk<cpx<T>>((T)(nr / den), (T)(ni / den));
  }
}

template <typename T>
__global__ void k_residual(const cpx<T>* __restrict__ kp1, const cpx<T>* __restrict__ k0,
                           const cpx<T>* __restrict__ p2, const cpx<T>* __restrict__ p3,
                           cpx<T>* __restrict__ out, long long n, int n2h1, KGeom g, Phys ph,
                           double a, double b) {
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    double k2, c2;
    natural_k(e, n2h1, g, ph, k2, c2);
    const cpx<T> s1 = kp1[e], s0 = k0[e], q2 = p2[e], q3 = p3[e];
    // sse_calculation.cu:133-140
    const double dr = -ph.Mres * k2 * (s0.x - a * q2.x + b * q3.x - c2 * s0.x);
    const double di = -ph.Mres * k2 * (s0.y - a * q2.y + b * q3.y - c2 * s0.y);
    out[e] = mk<cpx<T>>((T)((s1.x - s0.x) * ph.inv_dt - dr), (T)((s1.y - s0.y) * ph.inv_dt - di));
  }
}

// packed main-block offset of line l = a*n1 + b (plane stride xs)
__device__ __forceinline__ long long line_off(long long l, int n1, long long xs, int H) {
  const long long a = l / n1;
  return a * xs + (l - a * n1) * H;
}

template <typename T>
__global__ void k_pack(const cpx<T>* __restrict__ nat, cpx<T>* __restrict__ mainny, int n0n1, int H,
                       int n1, long long xs) {
  const long long tot = (long long)n0n1 * (H + 1);
  const long long nyoff = (long long)(n0n1 / n1) * xs;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
       e += (long long)gridDim.x * blockDim.x) {
    const int kz = (int)(e % (H + 1));
    const long long l = e / (H + 1);
    if (kz < H) mainny[line_off(l, n1, xs, H) + kz] = nat[e];
    else mainny[nyoff + l] = nat[e];
  }
}

template <typename T>
__global__ void k_unpack(const cpx<T>* __restrict__ mainny, cpx<T>* __restrict__ nat, int n0n1, int H,
                         int n1, long long xs) {
  const long long tot = (long long)n0n1 * (H + 1);
  const long long nyoff = (long long)(n0n1 / n1) * xs;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
       e += (long long)gridDim.x * blockDim.x) {
    const int kz = (int)(e % (H + 1));
    const long long l = e / (H + 1);
    nat[e] = kz < H ? mainny[line_off(l, n1, xs, H) + kz] : mainny[nyoff + l];
  }
}

template <typename S, typename D>
__global__ void k_convert(const S* __restrict__ src, D* __restrict__ dst, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    dst[i] = (D)src[i];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline int grid_for(long long n, int bs) {
  long long g = (n + bs - 1) / bs;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (int)g;
}

// ---------------------------------------------------------------------------
// noise initial condition on the device: the draws of xpfc_randn (Philox4x32-10
// with counter (cell_lo, cell_hi, 0, 0) and key (seed_lo, seed_hi), Box-Muller
// in FP64), so a job's field is built without a host round trip
// ---------------------------------------------------------------------------
__device__ __forceinline__ double randn_cell(unsigned long long cell, uint32_t k0, uint32_t k1) {
  uint32_t c0 = (uint32_t)cell, c1 = (uint32_t)(cell >> 32), c2 = 0u, c3 = 0u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
    const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
    const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  const double s32 = 1.0 / 4294967296.0;
  const double u0 = ((double)c0 + 0.5) * s32, u1 = ((double)c1 + 0.5) * s32;
  return sqrt(-2.0 * log(u0)) * cos(6.283185307179586 * u1);
}

// per-workgroup FP64 sums of the draws (fixed order)
__global__ __launch_bounds__(256) void k_noise_sum(long long n, uint32_t k0, uint32_t k1,
                                                   double* __restrict__ partial) {
  double a[1] = {0.0};
  for (long long c = (long long)blockIdx.x * 256 + threadIdx.x; c < n; c += (long long)gridDim.x * 256)
    a[0] += randn_cell((unsigned long long)c, k0, k1);
  block_sum<256, 1>(a, partial + blockIdx.x);
}

// amp * (z_c - mean_z), the fluctuation about the mean psi0, as reals at
// (c / plane) * pitch + c % plane
template <typename T>
__global__ __launch_bounds__(256) void k_noise_fill(T* __restrict__ out, long long n, long long plane,
                                                    long long pitch, uint32_t k0, uint32_t k1,
                                                    double amp, double mean_z) {
  for (long long c = (long long)blockIdx.x * 256 + threadIdx.x; c < n; c += (long long)gridDim.x * 256) {
    const long long a = c / plane;
    out[a * pitch + (c - a * plane)] = (T)(amp * (randn_cell((unsigned long long)c, k0, k1) - mean_z));
  }
}

int noise_partials(long long n) {
  return (int)std::min<long long>(NOISE_PARTIALS, std::max<long long>(1, (n + 255) / 256));
}

int launch_noise_sum(long long n, unsigned long long seed, double* partial, int* nblocks, hipStream_t s) {
  const int nb = noise_partials(n);
  *nblocks = nb;
  hipLaunchKernelGGL(k_noise_sum, dim3(nb), dim3(256), 0, s, n, (uint32_t)seed, (uint32_t)(seed >> 32),
                     partial);
  return launch_status();
}

int launch_noise_fill(int prec, void* out, long long n, long long plane, long long pitch,
                      unsigned long long seed, double amp, double mean_z, hipStream_t s) {
  const int nb = (int)std::min<long long>(8192, std::max<long long>(1, (n + 255) / 256));
  if (prec == 1)
    hipLaunchKernelGGL(k_noise_fill<float>, dim3(nb), dim3(256), 0, s, (float*)out, n, plane, pitch,
                       (uint32_t)seed, (uint32_t)(seed >> 32), amp, mean_z);
  else
    hipLaunchKernelGGL(k_noise_fill<double>, dim3(nb), dim3(256), 0, s, (double*)out, n, plane, pitch,
                       (uint32_t)seed, (uint32_t)(seed >> 32), amp, mean_z);
  return launch_status();
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free noise initialisation: the per-workgroup
// partial sums of k_noise_sum (at most NOISE_PARTIALS of them) live in this
// fixed device array instead of a context buffer.  Each GPU holds its own copy
// at its own address until the process ends; host code never passes the
// array's name to a kernel (on the host it is only a stand-in for the symbol)
// but the device address looked up once per device by noalloc_scratch_setup.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[NOISE_PARTIALS];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;                     // guards the table below
double* g_scratch_ptr[MAX_DEVICES] = {};     // device address of `scratch`, per device
std::mutex g_scratch_use[MAX_DEVICES];       // one user of a device's array at a time
}  // namespace

int noalloc_scratch_setup(int device) {
  if (device < 0 || device >= MAX_DEVICES) return -1;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return 0;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    return -2;
  }
  g_scratch_ptr[device] = static_cast<double*>(p);
  return 0;
}

double* noalloc_scratch(int device) {
  if (device < 0 || device >= MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[device];
}

std::mutex* noalloc_scratch_lock(int device) {
  return (device >= 0 && device < MAX_DEVICES) ? &g_scratch_use[device] : nullptr;
}

// Divergence probe: one workgroup; flag[s] = 1 when two modes of instance s
// (k = (0,0,1) and the first Nyquist-plane mode) are finite.  A non-finite
// value anywhere in real space reaches every mode at the next transform.
template <typename T>
__global__ void k_probe(const cpx<T>* __restrict__ psi, long long stride, long long iny, int K,
                        int* __restrict__ flag) {
  for (int s = threadIdx.x; s < K; s += blockDim.x) {
    const cpx<T> a = psi[(long long)s * stride + 1];
    const cpx<T> b = psi[(long long)s * stride + iny];
    flag[s] = (isfinite(a.x) && isfinite(a.y) && isfinite(b.x) && isfinite(b.y)) ? 1 : 0;
  }
}

int launch_probe(int prec, const void* psi, long long stride, long long iny, int K, int* flag,
                 hipStream_t s) {
  if (prec == 1)
    hipLaunchKernelGGL(k_probe<float>, dim3(1), dim3(256), 0, s, (const cpx<float>*)psi, stride,
                       iny, K, flag);
  else
    hipLaunchKernelGGL(k_probe<double>, dim3(1), dim3(256), 0, s, (const cpx<double>*)psi, stride,
                       iny, K, flag);
  return launch_status();
}

int launch_power(int prec, const void* src, void* dst, long long n, int p, bool f32, hipStream_t s) {
  if (prec == 1)
    hipLaunchKernelGGL(k_power<float>, dim3(grid_for(n, 256)), dim3(256), 0, s, (const float*)src, (float*)dst, n, p,
                       f32 ? 1 : 0);
  else
    hipLaunchKernelGGL(k_power<double>, dim3(grid_for(n, 256)), dim3(256), 0, s, (const double*)src, (double*)dst, n,
                       p, f32 ? 1 : 0);
  return launch_status();
}

int launch_scale(int prec, void* buf, long long n, double f, hipStream_t s) {
  if (prec == 1)
    hipLaunchKernelGGL(k_scale<float>, dim3(grid_for(n, 256)), dim3(256), 0, s, (float*)buf, n, (float)f);
  else
    hipLaunchKernelGGL(k_scale<double>, dim3(grid_for(n, 256)), dim3(256), 0, s, (double*)buf, n, f);
  return launch_status();
}

static long long natural_count(const KGeom& g, int n2) {
  return (long long)g.n0 * g.n1 * (n2 / 2 + 1);
}

int launch_convolve(int prec, const void* psi_k, void* out, const KGeom& g, const Phys& ph, int n2,
                    hipStream_t s) {
  const long long n = natural_count(g, n2);
  if (prec == 1)
    hipLaunchKernelGGL(k_convolve<float>, dim3(grid_for(n, 256)), dim3(256), 0, s,
                       (const cpx<float>*)psi_k, (cpx<float>*)out, n, n2 / 2 + 1, g, ph);
  else
    hipLaunchKernelGGL(k_convolve<double>, dim3(grid_for(n, 256)), dim3(256), 0, s,
                       (const cpx<double>*)psi_k, (cpx<double>*)out, n, n2 / 2 + 1, g, ph);
  return launch_status();
}

int launch_euler(int prec, const void* psi_k, const void* p2_k, const void* p3_k, void* out,
                 const KGeom& g, const Phys& ph, double a, double b, int n2, hipStream_t s) {
  const long long n = natural_count(g, n2);
  if (prec == 1)
    hipLaunchKernelGGL(k_euler<float>, dim3(grid_for(n, 256)), dim3(256), 0, s,
                       (const cpx<float>*)psi_k, (const cpx<float>*)p2_k, (const cpx<float>*)p3_k,
                       (cpx<float>*)out, n, n2 / 2 + 1, g, ph, a, b);
  else
    hipLaunchKernelGGL(k_euler<double>, dim3(grid_for(n, 256)), dim3(256), 0, s,
                       (const cpx<double>*)psi_k, (const cpx<double>*)p2_k, (const cpx<double>*)p3_k,
                       (cpx<double>*)out, n, n2 / 2 + 1, g, ph, a, b);
  return launch_status();
}

int launch_residual(int prec, const void* psi_kp1, const void* psi_k, const void* p2_k,
                    const void* p3_k, void* out, const KGeom& g, const Phys& ph, double a,
                    double b, int n2, hipStream_t s) {
  const long long n = natural_count(g, n2);
  if (prec == 1)
    hipLaunchKernelGGL(k_residual<float>, dim3(grid_for(n, 256)), dim3(256), 0, s,
                       (const cpx<float>*)psi_kp1, (const cpx<float>*)psi_k, (const cpx<float>*)p2_k,
                       (const cpx<float>*)p3_k, (cpx<float>*)out, n, n2 / 2 + 1, g, ph, a, b);
  else
    hipLaunchKernelGGL(k_residual<double>, dim3(grid_for(n, 256)), dim3(256), 0, s,
                       (const cpx<double>*)psi_kp1, (const cpx<double>*)psi_k, (const cpx<double>*)p2_k,
                       (const cpx<double>*)p3_k, (cpx<double>*)out, n, n2 / 2 + 1, g, ph, a, b);
  return launch_status();
}

int launch_pack(int prec, const void* natural, void* main_ny, int n0n1, int H, int n1, long long xs,
                hipStream_t s) {
  const long long n = (long long)n0n1 * (H + 1);
  if (prec == 1)
    hipLaunchKernelGGL(k_pack<float>, dim3(grid_for(n, 256)), dim3(256), 0, s,
                       (const cpx<float>*)natural, (cpx<float>*)main_ny, n0n1, H, n1, xs);
  else
    hipLaunchKernelGGL(k_pack<double>, dim3(grid_for(n, 256)), dim3(256), 0, s,
                       (const cpx<double>*)natural, (cpx<double>*)main_ny, n0n1, H, n1, xs);
  return launch_status();
}

int launch_unpack(int prec, const void* main_ny, void* natural, int n0n1, int H, int n1, long long xs,
                  hipStream_t s) {
  const long long n = (long long)n0n1 * (H + 1);
  if (prec == 1)
    hipLaunchKernelGGL(k_unpack<float>, dim3(grid_for(n, 256)), dim3(256), 0, s,
                       (const cpx<float>*)main_ny, (cpx<float>*)natural, n0n1, H, n1, xs);
  else
    hipLaunchKernelGGL(k_unpack<double>, dim3(grid_for(n, 256)), dim3(256), 0, s,
                       (const cpx<double>*)main_ny, (cpx<double>*)natural, n0n1, H, n1, xs);
  return launch_status();
}

int launch_to_double(int prec, const void* src, double* dst, long long n, hipStream_t s) {
  if (prec == 1)
    hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for(n, 256)), dim3(256), 0, s, (const float*)src, dst, n);
  else
    hipLaunchKernelGGL((k_convert<double, double>), dim3(grid_for(n, 256)), dim3(256), 0, s, (const double*)src, dst, n);
  return launch_status();
}

int launch_from_double(int prec, const double* src, void* dst, long long n, hipStream_t s) {
  if (prec == 1)
    hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(n, 256)), dim3(256), 0, s, src, (float*)dst, n);
  else
    hipLaunchKernelGGL((k_convert<double, double>), dim3(grid_for(n, 256)), dim3(256), 0, s, src, (double*)dst, n);
  return launch_status();
}

// ---------------------------------------------------------------------------
// plain streaming copy: the HBM rate of a kernel that only moves data
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_stream_copy(const float4* __restrict__ src, float4* __restrict__ dst,
                                                     long long n) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    dst[i] = src[i];
}

int launch_stream_copy(const void* src, void* dst, size_t bytes, hipStream_t s) {
  const long long n = (long long)(bytes / sizeof(float4));
  if (n <= 0) return 0;
  const long long nb = std::min<long long>((n + 255) / 256, 32768);
  hipLaunchKernelGGL(k_stream_copy, dim3((unsigned)nb), dim3(256), 0, s, (const float4*)src, (float4*)dst, n);
  return launch_status();
}

}  // namespace xpfc
