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

// ---- radices with a factor 3 or 5 ------------------------------------------
// cos and sin of d degrees for the angles of the radix-3/5/6/10/12/20/24
// butterflies (multiples of 15 and 18 degrees), as exact decimal constants.
struct CS {
  double c, s;
};
__host__ __device__ constexpr CS cs_first_quadrant(int d) {
  return d == 0    ? CS{1.0, 0.0}
         : d == 15 ? CS{0.96592582628906828674974320, 0.25881904510252076234889884}
         : d == 18 ? CS{0.95105651629515357211643933, 0.30901699437494742410229341}
         : d == 30 ? CS{0.86602540378443864676372317, 0.5}
         : d == 36 ? CS{0.80901699437494742410229341, 0.58778525229247312916870595}
         : d == 45 ? CS{0.70710678118654752440084436, 0.70710678118654752440084436}
         : d == 54 ? CS{0.58778525229247312916870595, 0.80901699437494742410229341}
         : d == 60 ? CS{0.5, 0.86602540378443864676372317}
         : d == 72 ? CS{0.30901699437494742410229341, 0.95105651629515357211643933}
         : d == 75 ? CS{0.25881904510252076234889884, 0.96592582628906828674974320}
         : d == 90 ? CS{0.0, 1.0}
                   : CS{0.0, 0.0};  // not an angle of these radices
}
__host__ __device__ constexpr CS cs_deg(int d) {
  d %= 360;
  return d <= 90    ? cs_first_quadrant(d)
         : d <= 180 ? CS{-cs_first_quadrant(180 - d).c, cs_first_quadrant(180 - d).s}
         : d <= 270 ? CS{-cs_first_quadrant(d - 180).c, -cs_first_quadrant(d - 180).s}
                    : CS{cs_first_quadrant(360 - d).c, -cs_first_quadrant(360 - d).s};
}

// a * exp(DIR * 2*pi*i * m / R) with compile-time m, R (R | 360)
template <int DIR, int R, int M, typename C> __device__ __forceinline__ C twiddle_const(C a) {
  using T = decltype(C::x);
  constexpr int d = (360 / R) * (M % R);
  if constexpr (d == 0) return a;
  else if constexpr (d == 90) return rot4<DIR>(a);
  else if constexpr (d == 180) return mk<C>(-a.x, -a.y);
  else if constexpr (d == 270) return rot4<-DIR>(a);
  else {
    constexpr CS w = cs_deg(d);
    return cmul(a, mk<C>((T)w.c, (T)(DIR * w.s)));
  }
}

template <int DIR, typename C> struct Dft<3, DIR, C> {
  static __device__ __forceinline__ void run(C* x) {
    using T = decltype(C::x);
    const T s = (T)(DIR * 0.86602540378443864676372317);  // DIR * sin(2 pi / 3)
    const C t = cadd(x[1], x[2]);
    const C d = csub(x[1], x[2]);
    const C m = mk<C>(x[0].x - (T)0.5 * t.x, x[0].y - (T)0.5 * t.y);  // cos(2 pi / 3) = -1/2
    const C r = mk<C>(-s * d.y, s * d.x);                               // DIR i sin(2 pi / 3) (x1 - x2)
    x[0] = cadd(x[0], t);
    x[1] = cadd(m, r);
    x[2] = csub(m, r);
  }
};
template <int DIR, typename C> struct Dft<5, DIR, C> {
  static __device__ __forceinline__ void run(C* x) {
    using T = decltype(C::x);
    const T c1 = (T)0.30901699437494742410229341, c2 = (T)-0.80901699437494742410229341;  // cos(2pi/5), cos(4pi/5)
    const T s1 = (T)(DIR * 0.95105651629515357211643933), s2 = (T)(DIR * 0.58778525229247312916870595);
    const C a1 = cadd(x[1], x[4]), b1 = csub(x[1], x[4]);
    const C a2 = cadd(x[2], x[3]), b2 = csub(x[2], x[3]);
    const C m1 = mk<C>(x[0].x + c1 * a1.x + c2 * a2.x, x[0].y + c1 * a1.y + c2 * a2.y);
    const C m2 = mk<C>(x[0].x + c2 * a1.x + c1 * a2.x, x[0].y + c2 * a1.y + c1 * a2.y);
    // i * (s1 b1 + s2 b2) and i * (s2 b1 - s1 b2)
    const C r1 = mk<C>(-(s1 * b1.y + s2 * b2.y), s1 * b1.x + s2 * b2.x);
    const C r2 = mk<C>(-(s2 * b1.y - s1 * b2.y), s2 * b1.x - s1 * b2.x);
    x[0] = mk<C>(x[0].x + a1.x + a2.x, x[0].y + a1.y + a2.y);
    x[1] = cadd(m1, r1);
    x[4] = csub(m1, r1);
    x[2] = cadd(m2, r2);
    x[3] = csub(m2, r2);
  }
};

// Composite R = P*Q (Cooley-Tukey, natural order in and out): Q-point DFTs of
// the P decimated sequences x[n1 + P n2], twiddles W_R^(n1 k2), then P-point
// DFTs giving X[k2 + Q k1].
template <int P, int Q, int DIR, typename C> struct DftPQ {
  template <int N1, int K2> static __device__ __forceinline__ void tw(C (&y)[P][Q]) {
    if constexpr (N1 < P) {
      if constexpr (K2 < Q) {
        y[N1][K2] = twiddle_const<DIR, P * Q, N1 * K2>(y[N1][K2]);
        tw<N1, K2 + 1>(y);
      } else {
        tw<N1 + 1, 0>(y);
      }
    }
  }
  static __device__ __forceinline__ void run(C* x) {
    C y[P][Q];
#pragma unroll
    for (int n1 = 0; n1 < P; ++n1) {
#pragma unroll
      for (int n2 = 0; n2 < Q; ++n2) y[n1][n2] = x[n1 + P * n2];
      Dft<Q, DIR, C>::run(y[n1]);
    }
    tw<1, 1>(y);
#pragma unroll
    for (int k2 = 0; k2 < Q; ++k2) {
      C z[P];
#pragma unroll
      for (int n1 = 0; n1 < P; ++n1) z[n1] = y[n1][k2];
      Dft<P, DIR, C>::run(z);
#pragma unroll
      for (int k1 = 0; k1 < P; ++k1) x[k2 + Q * k1] = z[k1];
    }
  }
};
template <int DIR, typename C> struct Dft<6, DIR, C> : DftPQ<2, 3, DIR, C> {};
template <int DIR, typename C> struct Dft<10, DIR, C> : DftPQ<2, 5, DIR, C> {};
template <int DIR, typename C> struct Dft<12, DIR, C> : DftPQ<4, 3, DIR, C> {};
template <int DIR, typename C> struct Dft<20, DIR, C> : DftPQ<4, 5, DIR, C> {};
template <int DIR, typename C> struct Dft<24, DIR, C> : DftPQ<8, 3, DIR, C> {};

// ---- stage schedule ----------------------------------------------------------
// Each stage takes the largest radix that divides both E (a thread's points,
// so a thread holds whole butterflies) and the length still to transform.
// For power-of-two L and E this is radix E greedily with the remainder last.
__host__ __device__ constexpr int pick_radix(int rem, int E) {
  for (int r = E; r > 1; --r)
    if (E % r == 0 && rem % r == 0) return r;
  return rem;
}
__host__ __device__ constexpr int radix_at(int L, int E, int s) {
  int rem = L;
  for (int i = 0; i < s; ++i) rem /= pick_radix(rem, E);
  return pick_radix(rem, E);
}
__host__ __device__ constexpr int nstages(int L, int E) {
  int rem = L, n = 0;
  while (rem > 1) { rem /= pick_radix(rem, E); ++n; }
  return n;
}
__host__ __device__ constexpr int ns_at(int L, int E, int s) {
  int ns = 1;
  for (int i = 0; i < s; ++i) ns *= radix_at(L, E, i);
  return ns;
}

// ---- points per thread -------------------------------------------------------
// Strided lines of L points: 8 for 2^k (radix-8 stages), 12 for 3*2^k (a
// radix-12 stage, then 4s and a 2), 20 for 5*2^k (radix 20, then 4s and a 2);
// a shorter line is one thread's.
template <int L> constexpr int elems_per_thread() {
  return L % 5 == 0 ? (L < 20 ? L : 20) : L % 3 == 0 ? (L < 12 ? L : 12) : (L < 8 ? L : 8);
}
// Real-space half lines of H points: for 2^k, radix-16 FP32 lines from H = 128
// up (one LDS exchange per transform at H = 128, 256) and 16 points per thread
// at H = 1024 (the line then fills one wavefront); for 3*2^k 12 points (24 in
// FP32 from H = 96: 384 = 24*8*2 instead of 12*4*4*2), 20 for 5*2^k.  Every
// choice leaves a power-of-two number of threads per line.
template <typename T, int H> constexpr int z_elems() {
  return H % 5 == 0   ? (H < 20 ? H : 20)
         : H % 3 == 0 ? (H < 12 ? H : (H >= 96 && sizeof(T) == 4) ? 24 : 12)
         : H < 8      ? H
         : (H >= 1024 || (H >= 128 && sizeof(T) == 4)) ? 16
                                                        : 8;
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
      const int jj = j % NS;
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
      const int jj = j % NS;
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
