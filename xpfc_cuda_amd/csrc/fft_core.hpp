// Device building blocks for the LDS-staged Stockham FFT passes (gfx950).
//
// A line of L complex points is spread over L/E threads; thread t holds the E
// points at rows t + m*(L/E), m = 0..E-1.  Every Stockham stage of radix R
// reads its butterfly j from rows j + q*L/R, so stage inputs always arrive in
// that register pattern; the last stage's outputs land in the same pattern,
// which lets a forward transform feed an inverse one without touching LDS.
// Between stages the outputs go through LDS (one write + one read per point).
//
// Twiddles come from a host-built table tw[m] = exp(-2*pi*i*m/L), computed
// in long double and stored in the working precision; the kernels copy it into
// LDS once per workgroup (stage_table), so a stage's twiddles are LDS reads
// (mostly broadcasts) instead of vector-memory loads.
#pragma once
#include <hip/hip_runtime.h>

namespace xpfc {

template <typename T> struct vec2;
template <> struct vec2<float> { using type = float2; };
template <> struct vec2<double> { using type = double2; };
template <typename T> using cpx = typename vec2<T>::type;

template <typename C> __device__ __forceinline__ C mk(decltype(C::x) x, decltype(C::x) y) {
  C r; r.x = x; r.y = y; return r;
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return mk<C>(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return mk<C>(a.x - b.x, a.y - b.y); }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  return mk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename C> __device__ __forceinline__ C cconj(C a) { return mk<C>(a.x, -a.y); }
template <typename C> __device__ __forceinline__ C cscale(C a, decltype(C::x) s) { return mk<C>(a.x * s, a.y * s); }

// multiply by exp(DIR * i*pi/2): DIR=-1 (forward) -> -i, DIR=+1 (inverse) -> +i
template <int DIR, typename C> __device__ __forceinline__ C rot4(C a) {
  if (DIR < 0) return mk<C>(a.y, -a.x);
  return mk<C>(-a.y, a.x);
}
// multiply by exp(DIR * i*pi/4)
template <int DIR, typename C> __device__ __forceinline__ C rot8_1(C a) {
  using T = decltype(C::x);
  const T c = (T)0.70710678118654752440;
  if (DIR < 0) return mk<C>(c * (a.x + a.y), c * (a.y - a.x));
  return mk<C>(c * (a.x - a.y), c * (a.x + a.y));
}
// multiply by exp(DIR * 3i*pi/4)
template <int DIR, typename C> __device__ __forceinline__ C rot8_3(C a) {
  using T = decltype(C::x);
  const T c = (T)0.70710678118654752440;
  if (DIR < 0) return mk<C>(c * (a.y - a.x), -c * (a.x + a.y));
  return mk<C>(-c * (a.x + a.y), c * (a.x - a.y));
}

// In-register R-point DFT, natural order in and out.
template <int R, int DIR, typename C> struct Dft;

template <int DIR, typename C> struct Dft<1, DIR, C> {
  static __device__ __forceinline__ void run(C*) {}
};
template <int DIR, typename C> struct Dft<2, DIR, C> {
  static __device__ __forceinline__ void run(C* x) {
    C a = x[0], b = x[1];
    x[0] = cadd(a, b); x[1] = csub(a, b);
  }
};
template <int DIR, typename C> struct Dft<4, DIR, C> {
  static __device__ __forceinline__ void run(C* x) {
    C a0 = cadd(x[0], x[2]), a1 = csub(x[0], x[2]);
    C a2 = cadd(x[1], x[3]), a3 = rot4<DIR>(csub(x[1], x[3]));
    x[0] = cadd(a0, a2); x[2] = csub(a0, a2);
    x[1] = cadd(a1, a3); x[3] = csub(a1, a3);
  }
};
template <int DIR, typename C> struct Dft<8, DIR, C> {
  static __device__ __forceinline__ void run(C* x) {
    C e[4] = {x[0], x[2], x[4], x[6]};
    C o[4] = {x[1], x[3], x[5], x[7]};
    Dft<4, DIR, C>::run(e);
    Dft<4, DIR, C>::run(o);
    o[1] = rot8_1<DIR>(o[1]);
    o[2] = rot4<DIR>(o[2]);
    o[3] = rot8_3<DIR>(o[3]);
#pragma unroll
    for (int q = 0; q < 4; ++q) { x[q] = cadd(e[q], o[q]); x[q + 4] = csub(e[q], o[q]); }
  }
};
template <int DIR, typename C> struct Dft<16, DIR, C> {
  static __device__ __forceinline__ void run(C* x) {
    using T = decltype(C::x);
    C e[8], o[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) { e[q] = x[2 * q]; o[q] = x[2 * q + 1]; }
    Dft<8, DIR, C>::run(e);
    Dft<8, DIR, C>::run(o);
    // o[q] *= exp(DIR * 2*pi*i*q/16)
    const T c1 = (T)0.92387953251128675613, s1 = (T)0.38268343236508977173;
    const T d = (T)DIR;
    o[1] = cmul(o[1], mk<C>(c1, d * s1));
    o[2] = rot8_1<DIR>(o[2]);
    o[3] = cmul(o[3], mk<C>(s1, d * c1));
    o[4] = rot4<DIR>(o[4]);
    o[5] = cmul(o[5], mk<C>(-s1, d * c1));
    o[6] = rot8_3<DIR>(o[6]);
    o[7] = cmul(o[7], mk<C>(-c1, d * s1));
#pragma unroll
    for (int q = 0; q < 8; ++q) { x[q] = cadd(e[q], o[q]); x[q + 8] = csub(e[q], o[q]); }
  }
};

// ---- stage schedule: greedy radix E, remainder last ------------------------
__host__ __device__ constexpr int radix_at(int L, int E, int s) {
  int rem = L;
  for (int i = 0; i < s; ++i) rem /= (rem >= E ? E : rem);
  return rem >= E ? E : rem;
}
__host__ __device__ constexpr int nstages(int L, int E) {
  int rem = L, n = 0;
  while (rem > 1) { rem /= (rem >= E ? E : rem); ++n; }
  return n;
}
__host__ __device__ constexpr int ns_at(int L, int E, int s) {
  int ns = 1;
  for (int i = 0; i < s; ++i) ns *= radix_at(L, E, i);
  return ns;
}

// One Stockham stage computed in registers (no data movement).
template <typename T, int L, int E, int R, int NS, int DIR>
__device__ __forceinline__ void stage_compute(cpx<T> (&v)[E], int t, const cpx<T>* __restrict__ tw) {
  using C = cpx<T>;
  constexpr int B = E / R;
#pragma unroll
  for (int u = 0; u < B; ++u) {
    C x[R];
#pragma unroll
    for (int q = 0; q < R; ++q) x[q] = v[u + q * B];
    if constexpr (NS > 1) {
      const int j = t + u * (L / E);
      const int jj = j & (NS - 1);
#pragma unroll
      for (int q = 1; q < R; ++q) {
        C w = tw[(q * jj) * (L / (NS * R))];
        if (DIR > 0) w.y = -w.y;
        x[q] = cmul(x[q], w);
      }
    }
    Dft<R, DIR, C>::run(x);
#pragma unroll
    for (int q = 0; q < R; ++q) v[u + q * B] = x[q];
  }
}

// Full L-point transform of the thread-distributed line (unnormalised).
// Idx maps a row index to this thread's LDS slot for that row.
template <typename T, int L, int E, int DIR, int S, class Idx>
__device__ __forceinline__ void fft_stages(cpx<T> (&v)[E], int t, cpx<T>* lds, const Idx& idx,
                                           const cpx<T>* __restrict__ tw) {
  constexpr int R = radix_at(L, E, S);
  constexpr int NS = ns_at(L, E, S);
  constexpr int NST = nstages(L, E);
  stage_compute<T, L, E, R, NS, DIR>(v, t, tw);
  if constexpr (S + 1 < NST) {
    constexpr int B = E / R;
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int j = t + u * (L / E);
      const int jj = j & (NS - 1);
      const int base = (j - jj) * R + jj;
#pragma unroll
      for (int q = 0; q < R; ++q) lds[idx(base + q * NS)] = v[u + q * B];
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = lds[idx(t + m * (L / E))];
    __syncthreads();
    fft_stages<T, L, E, DIR, S + 1>(v, t, lds, idx, tw);
  }
}

template <typename T, int L, int E, int DIR, class Idx>
__device__ __forceinline__ void fft_line(cpx<T> (&v)[E], int t, cpx<T>* lds, const Idx& idx,
                                         const cpx<T>* __restrict__ tw) {
  static_assert(L % E == 0, "E must divide L");
  if constexpr (L > 1) fft_stages<T, L, E, DIR, 0>(v, t, lds, idx, tw);
}

// Copy an n-entry table into LDS with all NT threads of the workgroup; the
// caller's next __syncthreads() publishes it.
template <int NT, typename C>
__device__ __forceinline__ void stage_table(C* dst, const C* __restrict__ src, int n) {
  for (int i = threadIdx.x; i < n; i += NT) dst[i] = src[i];
}

// Row-major tile of TC columns: row r, column c -> r*(TC+PAD) + c.
template <int TC, int PAD> struct TileIdx {
  int c;
  __device__ __forceinline__ int operator()(int r) const { return r * (TC + PAD) + c; }
};
// One line per group: line g, row r -> g*(LEN+PAD) + r.
template <int LEN, int PAD> struct LineIdx {
  int g;
  __device__ __forceinline__ int operator()(int r) const { return g * (LEN + PAD) + r; }
};

}  // namespace xpfc
