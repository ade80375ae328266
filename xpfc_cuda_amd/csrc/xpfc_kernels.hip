// HIP kernels of the XPFC time step for gfx950 (MI355X).
//
// Spectrum layout (packed half spectrum): for a real field of n0 x n1 x n2
// points (n2 contiguous; 2-D fields use n1 = 1) the spectrum is
//   main[n0][n1][H]  (kz = 0 .. H-1, H = n2/2)  followed by  ny[n0][n1] (kz = H).
// The k = 0 slot of `main` always holds 0: the mean density (the k = 0 mode)
// is carried separately in FP64 and added back in real space.
//
// One time step = four passes over HBM (three in 2-D):
//   k_xstep   : lines along axis 0: forward FFT of the nonlinear term, the
//               semi-implicit update of psi_k, inverse FFT      (euler_step.cu:50-59)
//   k_strided : lines along axis 1, inverse                      (3-D only)
//   k_zpass   : lines along axis 2: C2R, 1/N, N(psi) = -a psi^2 + b psi^3,
//               R2C                                               (power_functions.cu:34-53,
//                                                                  scaling_step.cu:48-49,
//                                                                  copy_matrix.cu:44-45)
//   k_strided : lines along axis 1, forward                      (3-D only)
// The sampled diagnostics (FE_calc.cu:64, sse_calculation.cu:133-140,
// Avg_calc.cu:3-18) are FP64 partial sums fused into k_xstep and k_zpass.
#include "fft_core.hpp"
#include "xpfc_internal.hpp"

namespace xpfc {

constexpr int PADT = 1;  // LDS padding (complex elements) per tile row
constexpr int PADZ = 1;  // LDS padding per line in the real-space pass (holds X[H])

// ---------------------------------------------------------------------------
// block reduction of K FP64 values, fixed order (deterministic)
// ---------------------------------------------------------------------------
template <int NT, int K>
__device__ __forceinline__ void block_sum(double (&a)[K], double* __restrict__ out) {
  constexpr int NW = (NT + 63) / 64;
  __shared__ double red[NW][K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
      if (off < NT) a[k] += __shfl_xor(a[k], off, 64);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) red[w][k] = a[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double s = 0.0;
      for (int i = 0; i < NW; ++i) s += red[i][k];
      out[k] = s;
    }
  }
}

// ---------------------------------------------------------------------------
// DCF: initialize.cu:209-217 (sum or max over peaks), FP64
// ---------------------------------------------------------------------------
__device__ __forceinline__ double c2hat(double kmag, const Phys& ph) {
  double r = 0.0;
  for (int i = 0; i < ph.npk; ++i) {
    const double d = kmag - ph.pk[i];
    const double e = d * d * ph.inv2a2[i];
    const double term = (e < 50.0) ? ph.amp[i] * exp(-e) : 0.0;
    r = ph.dcf_max ? fmax(r, term) : r + term;
  }
  return r;
}

__device__ __forceinline__ int signed_index(int i, int n) { return i <= n / 2 ? i : i - n; }

template <typename T> struct Cfg;  // strided-pass geometry per precision
template <> struct Cfg<float> {
  template <int L> static constexpr int TC = L >= 1024 ? 8 : 16;
};
template <> struct Cfg<double> {
  template <int L> static constexpr int TC = L >= 1024 ? 4 : 8;
};
template <int L> constexpr int elems_per_thread() { return L < 8 ? L : 8; }

// ---------------------------------------------------------------------------
// strided pass: forward (DIR=-1) or inverse (DIR=+1) transform of every line
// ---------------------------------------------------------------------------
template <typename T, int L, int DIR>
__global__ __launch_bounds__(Cfg<T>::template TC<L> * (L / elems_per_thread<L>()))
void k_strided(cpx<T>* __restrict__ buf, Segs segs, const cpx<T>* __restrict__ tw) {
  using C = cpx<T>;
  constexpr int TC = Cfg<T>::template TC<L>;
  constexpr int E = elems_per_thread<L>();
  __shared__ C lds[L * (TC + PADT)];
  const int tid = threadIdx.x;
  const int ci = tid % TC, t = tid / TC;
  int tile = blockIdx.x;
  const bool second = tile >= segs.tiles0;
  if (second) tile -= segs.tiles0;
  const Seg& sg = second ? segs.s[1] : segs.s[0];
  const int b = tile / sg.ntc, ct = tile - b * sg.ntc;
  const int col = ct * TC + ci;
  const bool act = col < sg.ncols;
  C* p = buf + sg.off + (long long)b * sg.bs + (long long)col * sg.cs;
  C v[E];
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = act ? p[(t + m * (L / E)) * sg.rs] : mk<C>(0, 0);
  fft_line<T, L, E, DIR>(v, t, lds, TileIdx<TC, PADT>{ci}, tw);
  if (act) {
#pragma unroll
    for (int m = 0; m < E; ++m) p[(t + m * (L / E)) * sg.rs] = v[m];
  }
}

// ---------------------------------------------------------------------------
// fused axis-0 pass: FFT(N) along axis 0, semi-implicit update, inverse FFT
// ---------------------------------------------------------------------------
template <typename T, int L, bool SAMPLE>
__global__ __launch_bounds__(Cfg<T>::template TC<L> * (L / elems_per_thread<L>()))
void k_xstep(cpx<T>* __restrict__ work, cpx<T>* __restrict__ psi, Segs segs,
             const cpx<T>* __restrict__ tw, KGeom g, Phys ph, double* __restrict__ partial) {
  using C = cpx<T>;
  constexpr int TC = Cfg<T>::template TC<L>;
  constexpr int E = elems_per_thread<L>();
  constexpr int NT = TC * (L / E);
  __shared__ C lds[L * (TC + PADT)];
  const int tid = threadIdx.x;
  const int ci = tid % TC, t = tid / TC;
  int tile = blockIdx.x;
  const bool second = tile >= segs.tiles0;
  if (second) tile -= segs.tiles0;
  const Seg& sg = second ? segs.s[1] : segs.s[0];
  const int b = tile / sg.ntc, ct = tile - b * sg.ntc;
  const int col = ct * TC + ci;
  const bool act = col < sg.ncols;
  const long long off = sg.off + (long long)b * sg.bs + (long long)col * sg.cs;
  C* pw = work + off;
  C* pp = psi + off;

  C v[E];
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = act ? pw[(t + m * (L / E)) * sg.rs] : mk<C>(0, 0);
  fft_line<T, L, E, -1>(v, t, lds, TileIdx<TC, PADT>{ci}, tw);

  // wavenumber of this column
  int kb, kz;
  if (sg.kind == 0) { kb = col / g.H; kz = col - kb * g.H; }
  else { kb = col; kz = g.H; }
  const double ky = g.dk1 * (double)signed_index(kb, g.n1);
  const double kzv = g.dk2 * (double)kz;
  const double wgt = act ? ((sg.kind == 0 && kz == 0) || sg.kind == 1 ? 1.0 : 2.0) : 0.0;
  double acc[XPART] = {0.0, 0.0};

#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int ka = t + m * (L / E);
    const double kx = g.dk0 * (double)signed_index(ka, g.n0);
    const double k2 = kx * kx + ky * ky + kzv * kzv;
    const double c2 = (k2 < ph.kcut2) ? c2hat(sqrt(k2), ph) : 0.0;
    // euler_step.cu:54 denominator, FP64
    const double den = 1.0 + k2 * ph.dtM * ph.one_m_eps - k2 * ph.dtM * c2;
    const double inv = 1.0 / den;
    const double coef = k2 * ph.dtM * inv;
    const C old = act ? pp[ka * sg.rs] : mk<C>(0, 0);
    const C nh = v[m];
    const double nx = inv * (double)old.x - coef * (double)nh.x;
    const double ny = inv * (double)old.y - coef * (double)nh.y;
    if constexpr (SAMPLE) {
      acc[0] += wgt * c2 * (nx * nx + ny * ny);
      // sse_calculation.cu:133-140 (eps absent from the residual, as there)
      const double rx = (nx - (double)old.x) * ph.inv_dt
                        + ph.Mres * k2 * ((double)old.x + (double)nh.x - c2 * (double)old.x);
      const double ry = (ny - (double)old.y) * ph.inv_dt
                        + ph.Mres * k2 * ((double)old.y + (double)nh.y - c2 * (double)old.y);
      acc[1] += wgt * (rx * rx + ry * ry);
    }
    const C nv = mk<C>((T)nx, (T)ny);
    if (act) pp[ka * sg.rs] = nv;
    v[m] = nv;
  }

  fft_line<T, L, E, +1>(v, t, lds, TileIdx<TC, PADT>{ci}, tw);
  if (act) {
#pragma unroll
    for (int m = 0; m < E; ++m) pw[(t + m * (L / E)) * sg.rs] = v[m];
  }
  if constexpr (SAMPLE) block_sum<NT, XPART>(acc, partial + (long long)blockIdx.x * XPART);
}

// ---------------------------------------------------------------------------
// real-space pass along the contiguous axis (H-point complex FFT of the
// even/odd packed line, with the half-spectrum pre/post twiddle)
// ---------------------------------------------------------------------------
template <int H> constexpr int z_elems() { return H < 8 ? H : 8; }
template <int H> constexpr int z_lines() {
  return (256 / (H / z_elems<H>())) > 0 ? 256 / (H / z_elems<H>()) : 1;
}

template <typename T, int H, int MODE, bool SAMPLE>
__global__ __launch_bounds__(z_lines<H>() * (H / z_elems<H>()))
void k_zpass(const void* __restrict__ in, const cpx<T>* __restrict__ in_ny, void* __restrict__ out,
             cpx<T>* __restrict__ out_ny, long long nlines, ZArgs za,
             const cpx<T>* __restrict__ tw, const cpx<T>* __restrict__ twp,
             double* __restrict__ partial) {
  using C = cpx<T>;
  constexpr int E = z_elems<H>();
  constexpr int TPL = H / E;  // threads per line
  constexpr int G = z_lines<H>();
  constexpr int NT = G * TPL;
  __shared__ C lds[G * (H + PADZ)];
  const int tid = threadIdx.x;
  const int g = tid / TPL, t = tid - g * TPL;
  const long long line = (long long)blockIdx.x * G + g;
  const bool act = line < nlines;
  const LineIdx<H, PADZ> idx{g};
  double acc[ZPART] = {0.0, 0.0, 0.0, 0.0};
  C v[E];

  if constexpr (MODE == Z_STEP || MODE == Z_C2R) {
    const C* sp = reinterpret_cast<const C*>(in) + line * H;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int k = t + m * TPL;
      lds[idx(k)] = act ? sp[k] : mk<C>(0, 0);
    }
    if (t == 0) lds[idx(H)] = act ? in_ny[line] : mk<C>(0, 0);
    __syncthreads();
    const T s = (T)za.scale;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int k = t + m * TPL;
      C xk = lds[idx(k)];
      C xm = lds[idx(H - k)];
      if (k == 0) { xk.y = 0; xm.y = 0; }  // DC and Nyquist bins are real (C2R)
      const C e = mk<C>(xk.x + xm.x, xk.y - xm.y);
      const C d = mk<C>(xk.x - xm.x, xk.y + xm.y);
      const C o = cmul(d, cconj(twp[k]));
      v[m] = mk<C>((e.x - o.y) * s, (e.y + o.x) * s);
    }
    __syncthreads();
    fft_line<T, H, E, +1>(v, t, lds, idx, tw);
    if constexpr (MODE == Z_C2R) {
      T* rp = reinterpret_cast<T*>(out) + line * (2 * H);
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int k = t + m * TPL;
        const double p0 = za.mean + (double)v[m].x, p1 = za.mean + (double)v[m].y;
        if constexpr (SAMPLE) {
          if (act) {
            acc[0] += p0 + p1;
            acc[1] += p0 * p0 + p1 * p1;
            acc[2] += p0 * p0 * p0 + p1 * p1 * p1;
            acc[3] += (p0 * p0) * (p0 * p0) + (p1 * p1) * (p1 * p1);
          }
        }
        if (act) reinterpret_cast<C*>(rp)[k] = mk<C>((T)p0, (T)p1);
      }
      if constexpr (SAMPLE) block_sum<NT, ZPART>(acc, partial + (long long)blockIdx.x * ZPART);
      return;
    } else {
      const T a = (T)za.a, b = (T)za.b;
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const double p0 = za.mean + (double)v[m].x, p1 = za.mean + (double)v[m].y;
        if constexpr (SAMPLE) {
          if (act) {
            acc[0] += p0 + p1;
            acc[1] += p0 * p0 + p1 * p1;
            acc[2] += p0 * p0 * p0 + p1 * p1 * p1;
            acc[3] += (p0 * p0) * (p0 * p0) + (p1 * p1) * (p1 * p1);
          }
        }
        const T q0 = (T)p0, q1 = (T)p1;
        v[m] = mk<C>(q0 * q0 * (b * q0 - a), q1 * q1 * (b * q1 - a));
      }
    }
  } else {
    // input: the fluctuation psi - mean as real pairs (x[2k], x[2k+1])
    const C* rp = reinterpret_cast<const C*>(in) + line * H;
    const T a = (T)za.a, b = (T)za.b;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int k = t + m * TPL;
      C x = act ? rp[k] : mk<C>(0, 0);
      if constexpr (MODE == Z_R2C_NL) {
        const T q0 = (T)(za.mean + (double)x.x), q1 = (T)(za.mean + (double)x.y);
        x = mk<C>(q0 * q0 * (b * q0 - a), q1 * q1 * (b * q1 - a));
      }
      v[m] = act ? x : mk<C>(0, 0);
    }
  }

  // forward H-point transform and the real-to-half-spectrum post twiddle
  fft_line<T, H, E, -1>(v, t, lds, idx, tw);
  __syncthreads();
#pragma unroll
  for (int m = 0; m < E; ++m) lds[idx(t + m * TPL)] = v[m];
  __syncthreads();
  C* op = reinterpret_cast<C*>(out) + line * H;
  const T half = (T)0.5;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int k = t + m * TPL;
    const C zk = lds[idx(k)];
    const C zm = lds[idx((H - k) & (H - 1))];
    const C e = mk<C>(half * (zk.x + zm.x), half * (zk.y - zm.y));
    const C o = mk<C>(half * (zk.y + zm.y), -half * (zk.x - zm.x));
    if (act) {
      if (k == 0) {
        op[0] = mk<C>(e.x + o.x, 0);
        out_ny[line] = mk<C>(e.x - o.x, 0);
      } else {
        op[k] = cadd(e, cmul(twp[k], o));
      }
    }
  }
  if constexpr (SAMPLE) block_sum<NT, ZPART>(acc, partial + (long long)blockIdx.x * ZPART);
}

// ---------------------------------------------------------------------------
// final fixed-order reduction of the per-workgroup partials
// out = {sum c2|psi_k|^2, sum |R_k|^2, sum psi, sum psi^2, sum psi^3, sum psi^4}
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_finalize(const double* __restrict__ px, int nx,
                                                   const double* __restrict__ pz, int nz,
                                                   double* __restrict__ out) {
  double a[NSUMS] = {0, 0, 0, 0, 0, 0};
  for (int i = threadIdx.x; i < nx; i += 1024) {
    a[0] += px[(long long)i * XPART + 0];
    a[1] += px[(long long)i * XPART + 1];
  }
  for (int i = threadIdx.x; i < nz; i += 1024) {
#pragma unroll
    for (int k = 0; k < ZPART; ++k) a[2 + k] += pz[(long long)i * ZPART + k];
  }
  block_sum<1024, NSUMS>(a, out);
}

// k-space Parseval sum  sum_k w c2(k) |psi_k|^2  over the packed spectrum
template <typename T>
__global__ __launch_bounds__(256) void k_kspace_sums(const cpx<T>* __restrict__ psi, long long nmain,
                                                     long long nny, KGeom g, Phys ph,
                                                     double* __restrict__ partial) {
  double acc[XPART] = {0.0, 0.0};
  const long long tot = nmain + nny;
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < tot;
       e += (long long)gridDim.x * 256) {
    int ka, kb, kz;
    double w;
    if (e < nmain) {
      kz = (int)(e % g.H);
      const long long r = e / g.H;
      kb = (int)(r % g.n1);
      ka = (int)(r / g.n1);
      w = kz == 0 ? 1.0 : 2.0;
    } else {
      const long long r = e - nmain;
      kb = (int)(r % g.n1);
      ka = (int)(r / g.n1);
      kz = g.H;
      w = 1.0;
    }
    const double kx = g.dk0 * signed_index(ka, g.n0), ky = g.dk1 * signed_index(kb, g.n1),
                 kzv = g.dk2 * kz;
    const double k2 = kx * kx + ky * ky + kzv * kzv;
    const double c2 = (k2 < ph.kcut2) ? c2hat(sqrt(k2), ph) : 0.0;
    const cpx<T> v = psi[e];
    acc[0] += w * c2 * ((double)v.x * v.x + (double)v.y * v.y);
  }
  block_sum<256, XPART>(acc, partial + (long long)blockIdx.x * XPART);
}

// ---------------------------------------------------------------------------
// operator-level kernels on natural layouts
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_power(const T* __restrict__ src, T* __restrict__ dst, long long n, int p) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const T x = src[i];
    T r = (T)1;
    for (int q = 0; q < p; ++q) r *= x;
    dst[i] = r;
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

template <typename T>
__global__ void k_pack(const cpx<T>* __restrict__ nat, cpx<T>* __restrict__ mainny, int n0n1, int H) {
  const long long tot = (long long)n0n1 * (H + 1);
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
       e += (long long)gridDim.x * blockDim.x) {
    const int kz = (int)(e % (H + 1));
    const long long l = e / (H + 1);
    if (kz < H) mainny[l * H + kz] = nat[e];
    else mainny[(long long)n0n1 * H + l] = nat[e];
  }
}

template <typename T>
__global__ void k_unpack(const cpx<T>* __restrict__ mainny, cpx<T>* __restrict__ nat, int n0n1, int H) {
  const long long tot = (long long)n0n1 * (H + 1);
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
       e += (long long)gridDim.x * blockDim.x) {
    const int kz = (int)(e % (H + 1));
    const long long l = e / (H + 1);
    nat[e] = kz < H ? mainny[l * H + kz] : mainny[(long long)n0n1 * H + l];
  }
}

template <typename S, typename D>
__global__ void k_convert(const S* __restrict__ src, D* __restrict__ dst, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    dst[i] = (D)src[i];
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------
static inline int grid_for(long long n, int bs) {
  long long g = (n + bs - 1) / bs;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (int)g;
}

template <typename T, int L> static Segs fill_tiles(const Segs& in) {
  constexpr int TC = Cfg<T>::template TC<L>;
  Segs s = in;
  int tiles[2];
  for (int i = 0; i < 2; ++i) {
    s.s[i].ntc = (s.s[i].ncols + TC - 1) / TC;
    tiles[i] = s.s[i].ntc * s.s[i].nbatch;
    if (s.s[i].ncols <= 0 || s.s[i].nbatch <= 0) tiles[i] = 0;
  }
  s.tiles0 = tiles[0];
  s.tiles1 = tiles[1];
  return s;
}

#define XPFC_LSWITCH(L, ...)                               \
  switch (L) {                                             \
    case 4: { constexpr int LL = 4; __VA_ARGS__; } break;       \
    case 8: { constexpr int LL = 8; __VA_ARGS__; } break;       \
    case 16: { constexpr int LL = 16; __VA_ARGS__; } break;     \
    case 32: { constexpr int LL = 32; __VA_ARGS__; } break;     \
    case 64: { constexpr int LL = 64; __VA_ARGS__; } break;     \
    case 128: { constexpr int LL = 128; __VA_ARGS__; } break;   \
    case 256: { constexpr int LL = 256; __VA_ARGS__; } break;   \
    case 512: { constexpr int LL = 512; __VA_ARGS__; } break;   \
    case 1024: { constexpr int LL = 1024; __VA_ARGS__; } break; \
    default: return -1;                            \
  }

template <typename T>
static int strided_T(void* buf, const Segs& segs0, int L, int dir, const void* tw, hipStream_t s) {
  XPFC_LSWITCH(L, {
    Segs sg = fill_tiles<T, LL>(segs0);
    constexpr int NT = Cfg<T>::template TC<LL> * (LL / elems_per_thread<LL>());
    const int nb = sg.tiles0 + sg.tiles1;
    if (nb == 0) return 0;
    if (dir < 0)
      hipLaunchKernelGGL((k_strided<T, LL, -1>), dim3(nb), dim3(NT), 0, s, (cpx<T>*)buf, sg,
                         (const cpx<T>*)tw);
    else
      hipLaunchKernelGGL((k_strided<T, LL, +1>), dim3(nb), dim3(NT), 0, s, (cpx<T>*)buf, sg,
                         (const cpx<T>*)tw);
  });
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_strided(int prec, void* buf, const Segs& segs, int L, int dir, const void* tw,
                   hipStream_t s) {
  return prec == 1 ? strided_T<float>(buf, segs, L, dir, tw, s)
                   : strided_T<double>(buf, segs, L, dir, tw, s);
}

template <typename T>
static int tiles_T(int L, const Segs& segs0, Segs* out) {
  XPFC_LSWITCH(L, { *out = fill_tiles<T, LL>(segs0); });
  return 0;
}
int strided_tiles(int prec, int L, const Segs& segs, Segs* out) {
  return prec == 1 ? tiles_T<float>(L, segs, out) : tiles_T<double>(L, segs, out);
}

template <typename T>
static int xstep_T(void* work, void* psi, const Segs& segs0, int L, const void* tw, const KGeom& g,
                   const Phys& ph, double* partial, bool sample, hipStream_t s) {
  XPFC_LSWITCH(L, {
    Segs sg = fill_tiles<T, LL>(segs0);
    constexpr int NT = Cfg<T>::template TC<LL> * (LL / elems_per_thread<LL>());
    const int nb = sg.tiles0 + sg.tiles1;
    if (sample)
      hipLaunchKernelGGL((k_xstep<T, LL, true>), dim3(nb), dim3(NT), 0, s, (cpx<T>*)work,
                         (cpx<T>*)psi, sg, (const cpx<T>*)tw, g, ph, partial);
    else
      hipLaunchKernelGGL((k_xstep<T, LL, false>), dim3(nb), dim3(NT), 0, s, (cpx<T>*)work,
                         (cpx<T>*)psi, sg, (const cpx<T>*)tw, g, ph, partial);
  });
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_xstep(int prec, void* work, void* psi, const Segs& segs, int L, const void* tw,
                 const KGeom& g, const Phys& ph, double* partial, bool sample, hipStream_t s) {
  return prec == 1 ? xstep_T<float>(work, psi, segs, L, tw, g, ph, partial, sample, s)
                   : xstep_T<double>(work, psi, segs, L, tw, g, ph, partial, sample, s);
}

int zpass_blocks(int H, long long nlines) {
  int G = 0;
  XPFC_LSWITCH(H, { G = z_lines<LL>(); });
  return (int)((nlines + G - 1) / G);
}

template <typename T, int H, int MODE>
static void zlaunch(const void* in, const void* in_ny, void* out, void* out_ny, long long nlines,
                    const ZArgs& za, const void* tw, const void* twp, double* partial, bool sample,
                    hipStream_t s) {
  constexpr int G = z_lines<H>();
  constexpr int NT = G * (H / z_elems<H>());
  const int nb = (int)((nlines + G - 1) / G);
  if (sample)
    hipLaunchKernelGGL((k_zpass<T, H, MODE, true>), dim3(nb), dim3(NT), 0, s, in,
                       (const cpx<T>*)in_ny, out, (cpx<T>*)out_ny, nlines, za,
                       (const cpx<T>*)tw, (const cpx<T>*)twp, partial);
  else
    hipLaunchKernelGGL((k_zpass<T, H, MODE, false>), dim3(nb), dim3(NT), 0, s, in,
                       (const cpx<T>*)in_ny, out, (cpx<T>*)out_ny, nlines, za,
                       (const cpx<T>*)tw, (const cpx<T>*)twp, partial);
}

template <typename T>
static int zpass_T(int mode, const void* in, const void* in_ny, void* out, void* out_ny,
                   long long nlines, int H, const ZArgs& za, const void* tw, const void* twp,
                   double* partial, bool sample, hipStream_t s) {
  XPFC_LSWITCH(H, {
    switch (mode) {
      case Z_STEP: zlaunch<T, LL, Z_STEP>(in, in_ny, out, out_ny, nlines, za, tw, twp, partial, sample, s); break;
      case Z_R2C_FLUCT: zlaunch<T, LL, Z_R2C_FLUCT>(in, in_ny, out, out_ny, nlines, za, tw, twp, partial, false, s); break;
      case Z_R2C_NL: zlaunch<T, LL, Z_R2C_NL>(in, in_ny, out, out_ny, nlines, za, tw, twp, partial, false, s); break;
      case Z_C2R: zlaunch<T, LL, Z_C2R>(in, in_ny, out, out_ny, nlines, za, tw, twp, partial, sample, s); break;
      default: return -1;
    }
  });
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_zpass(int prec, int mode, const void* in, const void* in_ny, void* out, void* out_ny,
                 long long nlines, int H, const ZArgs& za, const void* tw, const void* twp,
                 double* partial, bool sample, hipStream_t s) {
  return prec == 1 ? zpass_T<float>(mode, in, in_ny, out, out_ny, nlines, H, za, tw, twp, partial, sample, s)
                   : zpass_T<double>(mode, in, in_ny, out, out_ny, nlines, H, za, tw, twp, partial, sample, s);
}

int launch_finalize(const double* px, int nx, const double* pz, int nz, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, s, px, nx, pz, nz, out);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_kspace_sums(int prec, const void* psi, long long nmain, long long nny, const KGeom& g,
                       const Phys& ph, double* partial, int* nblocks, hipStream_t s) {
  const int nb = grid_for(nmain + nny, 256) > 4096 ? 4096 : grid_for(nmain + nny, 256);
  *nblocks = nb;
  if (prec == 1)
    hipLaunchKernelGGL(k_kspace_sums<float>, dim3(nb), dim3(256), 0, s, (const cpx<float>*)psi,
                       nmain, nny, g, ph, partial);
  else
    hipLaunchKernelGGL(k_kspace_sums<double>, dim3(nb), dim3(256), 0, s, (const cpx<double>*)psi,
                       nmain, nny, g, ph, partial);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_power(int prec, const void* src, void* dst, long long n, int p, hipStream_t s) {
  if (prec == 1)
    hipLaunchKernelGGL(k_power<float>, dim3(grid_for(n, 256)), dim3(256), 0, s, (const float*)src, (float*)dst, n, p);
  else
    hipLaunchKernelGGL(k_power<double>, dim3(grid_for(n, 256)), dim3(256), 0, s, (const double*)src, (double*)dst, n, p);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_scale(int prec, void* buf, long long n, double f, hipStream_t s) {
  if (prec == 1)
    hipLaunchKernelGGL(k_scale<float>, dim3(grid_for(n, 256)), dim3(256), 0, s, (float*)buf, n, (float)f);
  else
    hipLaunchKernelGGL(k_scale<double>, dim3(grid_for(n, 256)), dim3(256), 0, s, (double*)buf, n, f);
  return hipGetLastError() == hipSuccess ? 0 : -2;
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
  return hipGetLastError() == hipSuccess ? 0 : -2;
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
  return hipGetLastError() == hipSuccess ? 0 : -2;
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
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_pack(int prec, const void* natural, void* main_ny, int n0n1, int H, hipStream_t s) {
  const long long n = (long long)n0n1 * (H + 1);
  if (prec == 1)
    hipLaunchKernelGGL(k_pack<float>, dim3(grid_for(n, 256)), dim3(256), 0, s,
                       (const cpx<float>*)natural, (cpx<float>*)main_ny, n0n1, H);
  else
    hipLaunchKernelGGL(k_pack<double>, dim3(grid_for(n, 256)), dim3(256), 0, s,
                       (const cpx<double>*)natural, (cpx<double>*)main_ny, n0n1, H);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_unpack(int prec, const void* main_ny, void* natural, int n0n1, int H, hipStream_t s) {
  const long long n = (long long)n0n1 * (H + 1);
  if (prec == 1)
    hipLaunchKernelGGL(k_unpack<float>, dim3(grid_for(n, 256)), dim3(256), 0, s,
                       (const cpx<float>*)main_ny, (cpx<float>*)natural, n0n1, H);
  else
    hipLaunchKernelGGL(k_unpack<double>, dim3(grid_for(n, 256)), dim3(256), 0, s,
                       (const cpx<double>*)main_ny, (cpx<double>*)natural, n0n1, H);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_to_double(int prec, const void* src, double* dst, long long n, hipStream_t s) {
  if (prec == 1)
    hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for(n, 256)), dim3(256), 0, s, (const float*)src, dst, n);
  else
    hipLaunchKernelGGL((k_convert<double, double>), dim3(grid_for(n, 256)), dim3(256), 0, s, (const double*)src, dst, n);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_from_double(int prec, const double* src, void* dst, long long n, hipStream_t s) {
  if (prec == 1)
    hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(n, 256)), dim3(256), 0, s, src, (float*)dst, n);
  else
    hipLaunchKernelGGL((k_convert<double, double>), dim3(grid_for(n, 256)), dim3(256), 0, s, src, (double*)dst, n);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace xpfc
