// Runs the device FFT stages of xpfc_cuda_amd/csrc/fft_core.hpp on the CPU:
// a line of L points is spread over L/E host threads exactly as over the
// threads of a workgroup (thread t holds rows t + m*L/E), the LDS is a shared
// array whose every index is bounds-checked, and __syncthreads is a
// std::barrier.  Built by tests/test_fft_host.py with g++ against the stand-in
// header tests/fft_emu/hip/hip_runtime.h; test code only.
#include <atomic>
#include <barrier>
#include <cmath>
#include <complex>
#include <thread>
#include <vector>

#include "hip/hip_runtime.h"

thread_local EmuDim3 threadIdx;
namespace {
std::barrier<>* g_bar = nullptr;
std::atomic<int> g_bad_index{0};
}  // namespace
void __syncthreads() { g_bar->arrive_and_wait(); }

#include "../xpfc_cuda_amd/csrc/fft_core.hpp"

using namespace xpfc;

namespace {

template <int L>
struct CheckIdx {
  int operator()(int r) const {
    if (r < 0 || r >= L) g_bad_index.fetch_add(1);
    return (r < 0 || r >= L) ? 0 : r;
  }
};

// One transform of `in` (L points) with E points per thread; returns the
// largest |error| relative to max |X| of the exact DFT in long double, or -1
// when an LDS index left the line.
template <typename T, int L, int E, int DIR>
double check_line(const std::vector<std::complex<long double>>& in, int flip) {
  using C = cpx<T>;
  static_assert(L % E == 0, "E must divide L");
  constexpr int NT = L / E;
  // twiddle table as the library builds it: exp(-2 pi i m / L) in long double
  std::vector<C> tw(L);
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for (int m = 0; m < L; ++m) {
    const long double a = two_pi * (long double)m / (long double)L;
    tw[m].x = (T)std::cos(a);
    tw[m].y = (T)(-std::sin(a));
  }
  std::vector<C> lds(L), out(L);
  std::barrier<> bar(NT);
  g_bar = &bar;
  g_bad_index = 0;
  std::vector<std::thread> th;
  for (int t = 0; t < NT; ++t) {
    th.emplace_back([&, t]() {
      threadIdx.x = (unsigned)t;
      C v[E];
      for (int m = 0; m < E; ++m) {
        v[m].x = (T)in[t + m * NT].real();
        v[m].y = (T)in[t + m * NT].imag();
      }
      fft_line<T, L, E, DIR>(v, t, lds.data(), CheckIdx<L>{}, tw.data());
      for (int m = 0; m < E; ++m) out[t + m * NT] = v[m];
    });
  }
  for (auto& x : th) x.join();
  g_bar = nullptr;
  if (g_bad_index.load() != 0) return -1.0;
  long double emax = 0, xmax = 0;
  for (int k = 0; k < L; ++k) {
    std::complex<long double> s = 0;
    for (int n = 0; n < L; ++n) {
      const long double a = (long double)(flip ? -DIR : DIR) * two_pi * (long double)(((long long)n * k) % L) / (long double)L;
      s += in[n] * std::complex<long double>(std::cos(a), std::sin(a));
    }
    const std::complex<long double> got((long double)out[k].x, (long double)out[k].y);
    emax = std::max(emax, std::abs(got - s));
    xmax = std::max(xmax, std::abs(s));
  }
  return (double)(emax / xmax);
}

template <typename T, int L, int E>
double check_dir(int dir, const std::vector<std::complex<long double>>& in, int flip) {
  return dir < 0 ? check_line<T, L, E, -1>(in, flip) : check_line<T, L, E, +1>(in, flip);
}

template <int L>
double check_len(int shape, int prec, int dir, const std::vector<std::complex<long double>>& in, int* e,
                 int flip) {
  // shape 0: the strided passes' points per thread; 1: the real-space pass's
  if (shape == 0) {
    constexpr int E = elems_per_thread<L>();
    *e = E;
    return prec == 1 ? check_dir<float, L, E>(dir, in, flip) : check_dir<double, L, E>(dir, in, flip);
  }
  if (prec == 1) {
    constexpr int E = z_elems<float, L>();
    *e = E;
    return check_dir<float, L, E>(dir, in, flip);
  }
  constexpr int E = z_elems<double, L>();
  *e = E;
  return check_dir<double, L, E>(dir, in, flip);
}

}  // namespace

// err: relative max error of one forward (dir = -1) or inverse (+1)
// transform of a seeded random line; e: points per thread used.  flip != 0
// compares with the DFT of the opposite sign instead (a check that must fail).
// Returns 0, 1 for an unknown length, 2 for an LDS index out of the line.
extern "C" int fft_check(int L, int shape, int prec, int dir, unsigned seed, int flip, double* err, int* e) {
  std::vector<std::complex<long double>> in(L);
  unsigned long long st = seed * 6364136223846793005ULL + 1442695040888963407ULL;
  auto u = [&]() {
    st = st * 6364136223846793005ULL + 1442695040888963407ULL;
    return (long double)((st >> 11) & ((1ULL << 53) - 1)) / (long double)(1ULL << 53) * 2.0L - 1.0L;
  };
  for (auto& x : in) x = std::complex<long double>(u(), u());
  double r = 0;
  switch (L) {
#define XPFC_C(LEN) \
  case LEN: r = check_len<LEN>(shape, prec, dir, in, e, flip); break;
    XPFC_C(2) XPFC_C(3) XPFC_C(4) XPFC_C(5) XPFC_C(6) XPFC_C(8) XPFC_C(10) XPFC_C(12) XPFC_C(16) XPFC_C(20)
    XPFC_C(24) XPFC_C(32) XPFC_C(40) XPFC_C(48) XPFC_C(64) XPFC_C(80) XPFC_C(96) XPFC_C(128) XPFC_C(160)
    XPFC_C(192) XPFC_C(256) XPFC_C(320) XPFC_C(384) XPFC_C(512) XPFC_C(640) XPFC_C(768) XPFC_C(1024)
    XPFC_C(1280) XPFC_C(1536) XPFC_C(2048)
#undef XPFC_C
    default: return 1;
  }
  if (r < 0) return 2;
  *err = r;
  return 0;
}
