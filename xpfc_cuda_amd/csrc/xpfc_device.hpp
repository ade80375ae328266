// Device helpers shared by the pass kernels and the operator kernels.
#pragma once
#include "fft_core.hpp"
#include "xpfc_internal.hpp"

namespace xpfc {

// DCF (initialize.cu:209-217): sum or max over peaks of
// amp_i * exp(-(|k| - k_i)^2 / (2 alpha_i^2)), amp_i = exp(-(sigma k_i)^2/(2 rho_i B_i)).
// Terms with exponent >= ph.ecut (negligible in the denominator's precision)
// are skipped.
__device__ __forceinline__ double c2hat(double kmag, const Phys& ph) {
  double r = 0.0;
  for (int i = 0; i < ph.npk; ++i) {
    const double d = kmag - ph.pk[i];
    const double e = d * d * ph.inv2a2[i];
    const double term = (e < ph.ecut) ? ph.amp[i] * exp(-e) : 0.0;
    r = ph.dcf_max ? fmax(r, term) : r + term;
  }
  return r;
}

// C2 at k^2 = dksq * m on a cubic grid, from the instance's table
__device__ __forceinline__ double c2_lookup(int m, const Phys& ph) {
  return m < ph.mcut ? ph.c2tab[m] : 0.0;
}

// The semi-implicit update's coefficients at one mode (euler_step.cu:50-59):
// a2 = dt M k^2 and inv = 1 / (1 + a2 (1 - eps) - a2 C2), in FP64 (v_rcp_f64
// and two Newton steps).
__device__ __forceinline__ void update_coefs(double k2, double c2, const Phys& ph, double& a2, double& inv) {
  a2 = k2 * ph.dtM;
  const double den = 1.0 + a2 * ph.one_m_eps - a2 * c2;  // euler_step.cu:54
  double r = __builtin_amdgcn_rcp(den);
  double e = fma(-den, r, 1.0);
  r = fma(r, e, r);
  e = fma(-den, r, 1.0);
  inv = fma(r, e, r);
}

// Complex loads and stores with the non-temporal (streaming) cache hint: the
// line is not kept in the caches for reuse.
template <typename C> __device__ __forceinline__ C load_nt(const C* p) {
  using S = decltype(p->x);
  typedef S V __attribute__((ext_vector_type(2)));
  const V v = __builtin_nontemporal_load(reinterpret_cast<const V*>(p));
  C r;
  r.x = v.x;
  r.y = v.y;
  return r;
}
template <typename C> __device__ __forceinline__ void store_nt(C* p, C c) {
  using S = decltype(p->x);
  typedef S V __attribute__((ext_vector_type(2)));
  V v;
  v.x = c.x;
  v.y = c.y;
  __builtin_nontemporal_store(v, reinterpret_cast<V*>(p));
}

// FFT_freq.cu:17-24 index -> signed wavenumber index (Nyquist positive)
__device__ __forceinline__ int signed_index(int i, int n) { return i <= n / 2 ? i : i - n; }

// Block reduction of K FP64 values in a fixed order (deterministic):
// wave64 butterfly with __shfl_xor, then one LDS slot per wave.
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

}  // namespace xpfc
