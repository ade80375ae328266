// Ragged launches: the instances of several grids (size classes) stepped by
// one launch per pass.  Workgroups [wg0, wg0 + nact * wgs) of a launch belong
// to class k of its table, instance-major; a workgroup finds its class by
// scanning the table, remaps its index within the class to an XCD-ordered
// tile and runs the same per-line body as the one-grid kernels, so each
// instance's field is bitwise the one it gets alone.
//
// One kernel per line-length family (FAM 1: 2^k, 3: 3*2^k, 5: 5*2^k points):
// the bodies of every length of a family share one register allocation, so
// families are kept apart (a 12- or 20-point body would otherwise raise the
// allocation of every power-of-two class).  A pass whose classes span several
// families is one launch per family.  Included by the family units
// xpfc_rag_*.hip, which instantiate RagOps for their family.
#pragma once
#include <algorithm>

#include "xpfc_kernels.hpp"

namespace xpfc {

// the line lengths a family's ragged kernels compile: strided (axis 0 / 1)
// lines of up to 1024 points, real-space half lines of up to 512
#define XPFC_RAG_LINES_1(X) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024)
#define XPFC_RAG_LINES_3(X) X(12) X(24) X(48) X(96) X(192) X(384) X(768)
#define XPFC_RAG_LINES_5(X) X(20) X(40) X(80) X(160) X(320) X(640)
#define XPFC_RAG_HALVES_1(X) X(4) X(8) X(16) X(32) X(64) X(128) X(256) X(512)
#define XPFC_RAG_HALVES_3(X) X(6) X(12) X(24) X(48) X(96) X(192) X(384)
#define XPFC_RAG_HALVES_5(X) X(10) X(20) X(40) X(80) X(160) X(320)

// Runs body(std::integral_constant<int, L>) for the family's length L == v.
template <int FAM, class F>
__device__ __forceinline__ void rag_lines(int v, F&& body) {
#define XPFC_RC(LEN) \
  case LEN: body(std::integral_constant<int, LEN>{}); break;
  if constexpr (FAM == 1) {
    switch (v) { XPFC_RAG_LINES_1(XPFC_RC) default: break; }
  } else if constexpr (FAM == 3) {
    switch (v) { XPFC_RAG_LINES_3(XPFC_RC) default: break; }
  } else {
    switch (v) { XPFC_RAG_LINES_5(XPFC_RC) default: break; }
  }
}
template <int FAM, class F>
__device__ __forceinline__ void rag_halves(int v, F&& body) {
  if constexpr (FAM == 1) {
    switch (v) { XPFC_RAG_HALVES_1(XPFC_RC) default: break; }
  } else if constexpr (FAM == 3) {
    switch (v) { XPFC_RAG_HALVES_3(XPFC_RC) default: break; }
  } else {
    switch (v) { XPFC_RAG_HALVES_5(XPFC_RC) default: break; }
  }
#undef XPFC_RC
}

// the host's tile arithmetic (rag_tile) uses the points per thread of the FFT
#define XPFC_SA(LEN) static_assert(line_elems(LEN) == elems_per_thread<LEN>(), "points per thread");
XPFC_RAG_LINES_1(XPFC_SA) XPFC_RAG_LINES_3(XPFC_SA) XPFC_RAG_LINES_5(XPFC_SA)
#undef XPFC_SA

__device__ __forceinline__ int rag_class(const RClass* __restrict__ cls, int ncls) {
  int k = 0;
  for (int i = 1; i < ncls; ++i)
    if ((int)blockIdx.x >= cls[i].wg0) k = i;
  return k;
}

// LUT: every class of the launch has a square / cubic box (C2 from the
// table); the other form evaluates C2 per mode and needs more registers, so
// the two never share a kernel.
template <typename T, int FAM, bool SAMPLE, bool LUT>
__global__ __launch_bounds__(rag_nt(FAM)) void k_xstep_r(const RClass* __restrict__ cls, int ncls, int sib) {
  using C = cpx<T>;
  extern __shared__ __align__(16) unsigned char rag_smem[];
  const RClass& c = cls[rag_class(cls, ncls)];
  const int n = c.nact * c.wgs;
  int local = (int)blockIdx.x - c.wg0;
  if (local >= n) return;  // padding up to the next class
  local = xcd_tile(local, n, c.lgc);
  const int ii = local / c.wgs, tile = local - ii * c.wgs;
  const int slot = c.bt.slots[ii];
  if (c.budget && sib >= c.budget[slot]) return;
  C* work = (C*)c.a + (long long)slot * c.bt.stride;
  C* psi = (C*)c.b + (long long)slot * c.bt.stride;
  double* part = c.partial + (long long)slot * c.bt.pstride;
  const Phys& ph = c.bt.ia[slot].ph;
  C* lds = reinterpret_cast<C*>(rag_smem);
  rag_lines<FAM>(c.L, [&](auto len) {
    constexpr int LL = decltype(len)::value;
    constexpr int TC = rag_tile(LL);
    if constexpr (tile_ok<T, LL, TC>())
      xstep_body<T, LL, TC, SAMPLE, LUT>(work, psi, c.segs, tile, (const C*)c.tw, c.g, ph, lds,
                                         lds + LL * (TC + PADT), part, tile);
  });
}

template <typename T, int FAM, int DIR>
__global__ __launch_bounds__(rag_nt(FAM)) void k_strided_r(const RClass* __restrict__ cls, int ncls, int sib) {
  using C = cpx<T>;
  extern __shared__ __align__(16) unsigned char rag_smem[];
  const RClass& c = cls[rag_class(cls, ncls)];
  const int n = c.nact * c.wgs;
  int local = (int)blockIdx.x - c.wg0;
  if (local >= n) return;  // padding up to the next class
  local = xcd_tile(local, n, c.lgc);
  const int ii = local / c.wgs, tile = local - ii * c.wgs;
  const int slot = c.bt.slots[ii];
  if (c.budget && sib >= c.budget[slot]) return;
  // the inverse pass reads the state and writes the working buffer; the
  // forward pass works in place
  const long long so = (long long)slot * c.bt.stride;
  C* buf = (C*)c.a + so;
  const C* in = DIR > 0 ? (const C*)c.b + so : buf;
  C* lds = reinterpret_cast<C*>(rag_smem);
  rag_lines<FAM>(c.L, [&](auto len) {
    constexpr int LL = decltype(len)::value;
    constexpr int TC = rag_tile(LL);
    if constexpr (tile_ok<T, LL, TC>())
      strided_body<T, LL, TC, DIR, false>(in, buf, c.segs, tile, (const C*)c.tw, lds, lds + LL * (TC + PADT));
  });
}

// No occupancy hint: the bodies of every line length share one register
// allocation, which a 4-waves/SIMD bound would make spill.
template <typename T, int FAM, bool SAMPLE, bool PW>
__global__ __launch_bounds__(RAG_NTZ) void k_zpass_r(const RClass* __restrict__ cls, int ncls, int sib) {
  using C = cpx<T>;
  extern __shared__ __align__(16) unsigned char rag_smem[];
  const RClass& c = cls[rag_class(cls, ncls)];
  const int local = (int)blockIdx.x - c.wg0;
  const int ii = local / c.wgs, wb = local - ii * c.wgs;
  if (ii >= c.nact) return;  // padding up to the next class
  const int slot = c.bt.slots[ii];
  if (c.budget && sib >= c.budget[slot]) return;
  const long long so = (long long)slot * c.bt.stride;
  C* base = (C*)c.a + so;
  const C* in = (const C*)c.b + so;
  double* part = c.partial + (long long)slot * c.bt.pzstride;
  C* lds = reinterpret_cast<C*>(rag_smem);
  rag_halves<FAM>(c.L, [&](auto half) {
    constexpr int HH = decltype(half)::value;
    if constexpr (z_ok<T, HH, RAG_NTZ>())
      zpass_body<T, HH, RAG_NTZ, Z_STEP, SAMPLE, PW>(in, in + c.nyoff, base, base + c.nyoff, c.nlines, c.zl,
                                                      c.bt.ia[slot].za, (const C*)c.tw, (const C*)c.twp, part,
                                                      lds, wb);
  });
}

namespace {
// dynamic LDS beyond the default limit needs the per-kernel attribute
template <class K>
bool allow_lds(K kernel, size_t bytes) {
  if (bytes <= 65536) return true;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)bytes) == hipSuccess;
}
}  // namespace

// the dynamic LDS of the family's largest strided tile and real-space block
template <typename T, int FAM>
size_t rag_max_lds(bool strided) {
  size_t m = 0;
#define XPFC_RC(LEN)                                                                         \
  if constexpr (tile_ok<T, LEN, rag_tile(LEN)>())                                            \
    m = std::max(m, ((size_t)LEN * (rag_tile(LEN) + PADT) + LEN) * sizeof(cpx<T>));
#define XPFC_RZ(HH) \
  if constexpr (z_ok<T, HH, RAG_NTZ>()) m = std::max(m, (size_t)z_lds_elems<T, HH, RAG_NTZ>() * sizeof(cpx<T>));
  if (strided) {
    if constexpr (FAM == 1) { XPFC_RAG_LINES_1(XPFC_RC) }
    else if constexpr (FAM == 3) { XPFC_RAG_LINES_3(XPFC_RC) }
    else { XPFC_RAG_LINES_5(XPFC_RC) }
  } else {
    if constexpr (FAM == 1) { XPFC_RAG_HALVES_1(XPFC_RZ) }
    else if constexpr (FAM == 3) { XPFC_RAG_HALVES_3(XPFC_RZ) }
    else { XPFC_RAG_HALVES_5(XPFC_RZ) }
  }
#undef XPFC_RC
#undef XPFC_RZ
  return m;
}

// The attribute is set to the family's maximum, never lowered, so class
// tables built earlier stay launchable.
template <typename T, int FAM>
int RagOps<T, FAM>::prepare() {
  const size_t lx = rag_max_lds<T, FAM>(true), lz = rag_max_lds<T, FAM>(false);
  const bool ok = allow_lds(k_xstep_r<T, FAM, true, true>, lx) && allow_lds(k_xstep_r<T, FAM, false, true>, lx) &&
                  allow_lds(k_xstep_r<T, FAM, true, false>, lx) && allow_lds(k_xstep_r<T, FAM, false, false>, lx) &&
                  allow_lds(k_strided_r<T, FAM, -1>, lx) && allow_lds(k_strided_r<T, FAM, 1>, lx) &&
                  allow_lds(k_zpass_r<T, FAM, true, true>, lz) && allow_lds(k_zpass_r<T, FAM, true, false>, lz) &&
                  allow_lds(k_zpass_r<T, FAM, false, true>, lz) && allow_lds(k_zpass_r<T, FAM, false, false>, lz);
  return ok ? 0 : launch_status(-2);
}

template <typename T, int FAM>
int RagOps<T, FAM>::xstep(const RClass* cls, int ncls, int nwg, size_t lds, bool lut, bool sample, int sib,
                          hipStream_t s) {
  if (nwg <= 0) return 0;
  const dim3 g(nwg), b(rag_nt(FAM));
  if (lut) {
    if (sample) hipLaunchKernelGGL((k_xstep_r<T, FAM, true, true>), g, b, lds, s, cls, ncls, sib);
    else hipLaunchKernelGGL((k_xstep_r<T, FAM, false, true>), g, b, lds, s, cls, ncls, sib);
  } else {
    if (sample) hipLaunchKernelGGL((k_xstep_r<T, FAM, true, false>), g, b, lds, s, cls, ncls, sib);
    else hipLaunchKernelGGL((k_xstep_r<T, FAM, false, false>), g, b, lds, s, cls, ncls, sib);
  }
  return launch_status();
}

template <typename T, int FAM>
int RagOps<T, FAM>::strided(const RClass* cls, int ncls, int nwg, size_t lds, int dir, int sib, hipStream_t s) {
  if (nwg <= 0) return 0;
  if (dir < 0)
    hipLaunchKernelGGL((k_strided_r<T, FAM, -1>), dim3(nwg), dim3(rag_nt(FAM)), lds, s, cls, ncls, sib);
  else
    hipLaunchKernelGGL((k_strided_r<T, FAM, 1>), dim3(nwg), dim3(rag_nt(FAM)), lds, s, cls, ncls, sib);
  return launch_status();
}

template <typename T, int FAM>
int RagOps<T, FAM>::zpass(const RClass* cls, int ncls, int nwg, size_t lds, bool sample, bool pw, int sib,
                          hipStream_t s) {
  if (nwg <= 0) return 0;
  const dim3 g(nwg), b(RAG_NTZ);
  if (sample) {
    if (pw) hipLaunchKernelGGL((k_zpass_r<T, FAM, true, true>), g, b, lds, s, cls, ncls, sib);
    else hipLaunchKernelGGL((k_zpass_r<T, FAM, true, false>), g, b, lds, s, cls, ncls, sib);
  } else {
    if (pw) hipLaunchKernelGGL((k_zpass_r<T, FAM, false, true>), g, b, lds, s, cls, ncls, sib);
    else hipLaunchKernelGGL((k_zpass_r<T, FAM, false, false>), g, b, lds, s, cls, ncls, sib);
  }
  return launch_status();
}

template <typename T, int FAM>
int RagOps<T, FAM>::strided_wgs(int L, const Segs& segs) {
#define XPFC_RC(LEN)                                         \
  case LEN: {                                                \
    constexpr int TC = rag_tile(LEN);                        \
    if constexpr (tile_ok<T, LEN, TC>()) {                   \
      const Segs sg = fill_tiles(segs, TC);                  \
      return sg.tiles0 + sg.tiles1;                          \
    }                                                        \
    return -3;                                               \
  }
  if constexpr (FAM == 1) {
    switch (L) { XPFC_RAG_LINES_1(XPFC_RC) default: return -1; }
  } else if constexpr (FAM == 3) {
    switch (L) { XPFC_RAG_LINES_3(XPFC_RC) default: return -1; }
  } else {
    switch (L) { XPFC_RAG_LINES_5(XPFC_RC) default: return -1; }
  }
#undef XPFC_RC
}

template <typename T, int FAM>
long long RagOps<T, FAM>::zpass_wgs(int H, long long nlines, size_t* lds) {
#define XPFC_RC(HH)                                                            \
  case HH: {                                                                   \
    if constexpr (z_ok<T, HH, RAG_NTZ>()) {                                    \
      constexpr int G = RAG_NTZ / z_tpl<T, HH>();                              \
      if (lds) *lds = (size_t)z_lds_elems<T, HH, RAG_NTZ>() * sizeof(cpx<T>); \
      return (nlines + G - 1) / G;                                             \
    }                                                                          \
    return -3;                                                                 \
  }
  if constexpr (FAM == 1) {
    switch (H) { XPFC_RAG_HALVES_1(XPFC_RC) default: return -1; }
  } else if constexpr (FAM == 3) {
    switch (H) { XPFC_RAG_HALVES_3(XPFC_RC) default: return -1; }
  } else {
    switch (H) { XPFC_RAG_HALVES_5(XPFC_RC) default: return -1; }
  }
#undef XPFC_RC
}

#define XPFC_INSTANTIATE_RAG(FAM)       \
  template struct RagOps<float, FAM>;   \
  template struct RagOps<double, FAM>;

}  // namespace xpfc
