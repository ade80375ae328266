// HIP kernels of the XPFC time step for gfx950 (MI355X): the dispatch of the
// step passes on the runtime line length (the per-length launchers are
// compiled in xpfc_len_*.hip), the fixed-order reductions, the C2 table, and
// the ragged launches that step instances of several grids in one launch per
// pass.  The pass kernels themselves are in xpfc_kernels.hpp.
#include <type_traits>

#include "xpfc_kernels.hpp"
#include "xpfc_lengths.hpp"

namespace xpfc {

// ---------------------------------------------------------------------------
// final fixed-order reduction of the per-workgroup partials
// out = {sum c2|psi_k|^2, sum |R_k|^2, sum psi, sum psi^2, sum psi^3, sum psi^4}
// ---------------------------------------------------------------------------
// one workgroup per listed instance: its partials -> out[slot * NSUMS ...]
__global__ __launch_bounds__(1024) void k_finalize(const double* __restrict__ px, int nx,
                                                   const double* __restrict__ pz, int nz,
                                                   Batch bt, double* __restrict__ out) {
  const int slot = bt.slots[blockIdx.x];
  px += (long long)slot * bt.pstride;
  pz += (long long)slot * bt.pzstride;
  out += (long long)slot * NSUMS;
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
                                                     long long nny, KGeom g, Batch bt,
                                                     double* __restrict__ partial) {
  const int slot = bt.slots[0];
  psi += (long long)slot * bt.stride;
  partial += (long long)slot * bt.pstride;
  const Phys& ph = bt.ia[slot].ph;
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
    const int ia = signed_index(ka, g.n0), jb = signed_index(kb, g.n1);
    double c2;
    if (g.cubic) {  // the same C2 values as the axis-0 pass
      c2 = c2_lookup(ia * ia + jb * jb + kz * kz, ph);
    } else {
      const double kx = g.dk0 * ia, ky = g.dk1 * jb, kzv = g.dk2 * kz;
      const double k2 = kx * kx + ky * ky + kzv * kzv;
      c2 = (k2 < ph.kcut2) ? c2hat(sqrt(k2), ph) : 0.0;
    }
    const cpx<T> v = psi[e < nmain ? (long long)ka * g.xs + (long long)kb * g.H + kz
                                  : g.nyoff + (e - nmain)];
    acc[0] += w * c2 * ((double)v.x * v.x + (double)v.y * v.y);
  }
  block_sum<256, XPART>(acc, partial + (long long)blockIdx.x * XPART);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
namespace {
thread_local hipError_t t_launch_error = hipSuccess;
}

int launch_status(int ok_value) {
  const hipError_t e = hipGetLastError();
  if (e == hipSuccess) return ok_value;
  t_launch_error = e;
  return -2;
}

hipError_t launch_error() { return t_launch_error; }

// Each dispatcher returns -1 for a length that is not compiled.
namespace {
template <typename T>
int strided_T(void* buf, const Segs& segs, int L, int tc, int dir, const void* tw, const Batch& bt, int ny,
              bool nt, hipStream_t s) {
  switch (L) {
#define XPFC_C(LEN) \
  case LEN: return LenOps<T, LEN>::strided(buf, segs, tc, dir, tw, bt, ny, nt, s);
    XPFC_LENGTHS(XPFC_C)
#undef XPFC_C
    default: return -1;
  }
}
template <typename T>
int blocks_T(int L, int tc, const Segs& segs) {
  switch (L) {
#define XPFC_C(LEN) \
  case LEN: return LenOps<T, LEN>::blocks(tc, segs);
    XPFC_LENGTHS(XPFC_C)
#undef XPFC_C
    default: return -1;
  }
}
template <typename T>
int xstep_T(void* work, void* psi, const Segs& segs, int L, int tc, const void* tw, const KGeom& g,
            const Batch& bt, int ny, double* partial, bool sample, hipStream_t s) {
  switch (L) {
#define XPFC_C(LEN) \
  case LEN: return LenOps<T, LEN>::xstep(work, psi, segs, tc, tw, g, bt, ny, partial, sample, s);
    XPFC_LENGTHS(XPFC_C)
#undef XPFC_C
    default: return -1;
  }
}
template <typename T>
int zpass_T(int mode, const void* in, const void* in_ny, void* out, void* out_ny, long long nlines, int H,
            const ZLines& zl, int ntz, const Batch& bt, int ny, const void* tw, const void* twp, double* partial,
            bool sample, bool pw, hipStream_t s) {
  switch (H) {
#define XPFC_C(LEN)                                                                                       \
  case LEN:                                                                                               \
    return LenOps<T, LEN>::zpass(mode, in, in_ny, out, out_ny, nlines, zl, ntz, bt, ny, tw, twp, partial, \
                                 sample, pw, s);
    XPFC_LENGTHS(XPFC_C)
#undef XPFC_C
    default: return -1;
  }
}
template <typename T>
int zblocks_T(int H, int ntz, long long nlines) {
  switch (H) {
#define XPFC_C(LEN) \
  case LEN: return LenOps<T, LEN>::zblocks(ntz, nlines);
    XPFC_LENGTHS(XPFC_C)
#undef XPFC_C
    default: return -1;
  }
}
}  // namespace

int launch_strided(int prec, void* buf, const Segs& segs, int L, int tc, int dir, const void* tw,
                   const Batch& bt, int ny, bool nt, hipStream_t s) {
  return prec == 1 ? strided_T<float>(buf, segs, L, tc, dir, tw, bt, ny, nt, s)
                   : strided_T<double>(buf, segs, L, tc, dir, tw, bt, ny, nt, s);
}

int strided_blocks(int prec, int L, int tc, const Segs& segs) {
  return prec == 1 ? blocks_T<float>(L, tc, segs) : blocks_T<double>(L, tc, segs);
}

// tc: tile width; the square / cubic table path is taken when g.cubic is set.
// Returns the workgroup count (= number of X partials).
int launch_xstep(int prec, void* work, void* psi, const Segs& segs, int L, int tc, const void* tw,
                 const KGeom& g, const Batch& bt, int ny, double* partial, bool sample,
                 hipStream_t s) {
  return prec == 1 ? xstep_T<float>(work, psi, segs, L, tc, tw, g, bt, ny, partial, sample, s)
                   : xstep_T<double>(work, psi, segs, L, tc, tw, g, bt, ny, partial, sample, s);
}

int zpass_blocks(int prec, int H, int ntz, long long nlines) {
  return prec == 1 ? zblocks_T<float>(H, ntz, nlines) : zblocks_T<double>(H, ntz, nlines);
}

int launch_zpass(int prec, int mode, const void* in, const void* in_ny, void* out, void* out_ny,
                 long long nlines, int H, const ZLines& zl, int ntz, const Batch& bt, int ny,
                 const void* tw, const void* twp, double* partial, bool sample, bool pw, hipStream_t s) {
  return prec == 1
             ? zpass_T<float>(mode, in, in_ny, out, out_ny, nlines, H, zl, ntz, bt, ny, tw, twp, partial, sample, pw, s)
             : zpass_T<double>(mode, in, in_ny, out, out_ny, nlines, H, zl, ntz, bt, ny, tw, twp, partial, sample, pw, s);
}

int launch_finalize(const double* px, int nx, const double* pz, int nz, const Batch& bt, int ny,
                    double* out, hipStream_t s) {
  if (ny == 0) return 0;
  hipLaunchKernelGGL(k_finalize, dim3(ny), dim3(1024), 0, s, px, nx, pz, nz, bt, out);
  return launch_status();
}

int launch_kspace_sums(int prec, const void* psi, long long nmain, long long nny, const KGeom& g,
                       const Batch& bt, double* partial, int* nblocks, hipStream_t s) {
  long long nb = (nmain + nny + 255) / 256;
  if (nb > 4096) nb = 4096;
  if (nb < 1) nb = 1;
  *nblocks = (int)nb;
  if (prec == 1)
    hipLaunchKernelGGL(k_kspace_sums<float>, dim3(nb), dim3(256), 0, s, (const cpx<float>*)psi,
                       nmain, nny, g, bt, partial);
  else
    hipLaunchKernelGGL(k_kspace_sums<double>, dim3(nb), dim3(256), 0, s, (const cpx<double>*)psi,
                       nmain, nny, g, bt, partial);
  return launch_status();
}

// C2 table of one instance over the integer m = |k|^2 / dksq
__global__ __launch_bounds__(256) void k_c2table(double* __restrict__ tab, int mcut, double dksq,
                                                 Phys ph) {
  const int m = blockIdx.x * 256 + threadIdx.x;
  if (m < mcut) tab[m] = c2hat(sqrt(dksq * (double)m), ph);
}

int launch_c2table(double* tab, int mcut, double dksq, const Phys& ph, hipStream_t s) {
  if (mcut <= 0) return 0;
  hipLaunchKernelGGL(k_c2table, dim3((mcut + 255) / 256), dim3(256), 0, s, tab, mcut, dksq, ph);
  return launch_status();
}

// ---------------------------------------------------------------------------
// ragged launches: several size classes in one launch per pass
// ---------------------------------------------------------------------------
__device__ __forceinline__ int rag_class(const RClass* __restrict__ cls, int ncls) {
  int k = 0;
  for (int i = 1; i < ncls; ++i)
    if ((int)blockIdx.x >= cls[i].wg0) k = i;
  return k;
}

#define XPFC_RLSWITCH(L, ...)                                    \
  switch (L) {                                                   \
    case 8: { constexpr int LL = 8; __VA_ARGS__; } break;        \
    case 16: { constexpr int LL = 16; __VA_ARGS__; } break;      \
    case 32: { constexpr int LL = 32; __VA_ARGS__; } break;      \
    case 64: { constexpr int LL = 64; __VA_ARGS__; } break;      \
    case 128: { constexpr int LL = 128; __VA_ARGS__; } break;    \
    case 256: { constexpr int LL = 256; __VA_ARGS__; } break;    \
    case 512: { constexpr int LL = 512; __VA_ARGS__; } break;    \
    case 1024: { constexpr int LL = 1024; __VA_ARGS__; } break;  \
    default: break;                                              \
  }

// LUT: every class of the launch has a square / cubic box (C2 from the
// table); the other form evaluates C2 per mode and needs more registers, so
// the two never share a kernel.
template <typename T, bool SAMPLE, bool LUT>
__global__ __launch_bounds__(RAG_NT) void k_xstep_r(const RClass* __restrict__ cls, int ncls, int sib) {
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
  XPFC_RLSWITCH(c.L, {
    constexpr int TC = rag_tile(LL);
    if constexpr (tile_ok<T, LL, TC>()) {
      xstep_body<T, LL, TC, SAMPLE, LUT>(work, psi, c.segs, tile, (const C*)c.tw, c.g, ph, lds,
                                         lds + LL * (TC + PADT), part, tile);
    }
  })
}

template <typename T, int DIR>
__global__ __launch_bounds__(RAG_NT) void k_strided_r(const RClass* __restrict__ cls, int ncls, int sib) {
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
  C* buf = (C*)c.a + (long long)slot * c.bt.stride;
  C* lds = reinterpret_cast<C*>(rag_smem);
  XPFC_RLSWITCH(c.L, {
    constexpr int TC = rag_tile(LL);
    if constexpr (tile_ok<T, LL, TC>())
      strided_body<T, LL, TC, DIR, false>(buf, c.segs, tile, (const C*)c.tw, lds, lds + LL * (TC + PADT));
  })
}

// No occupancy hint: the bodies of every line length share one register
// allocation, which a 4-waves/SIMD bound would make spill.
template <typename T, bool SAMPLE, bool PW>
__global__ __launch_bounds__(RAG_NTZ)
void k_zpass_r(const RClass* __restrict__ cls, int ncls, int sib) {
  using C = cpx<T>;
  extern __shared__ __align__(16) unsigned char rag_smem[];
  const RClass& c = cls[rag_class(cls, ncls)];
  const int local = (int)blockIdx.x - c.wg0;
  const int ii = local / c.wgs, wb = local - ii * c.wgs;
  if (ii >= c.nact) return;  // padding up to the next class
  const int slot = c.bt.slots[ii];
  if (c.budget && sib >= c.budget[slot]) return;
  C* base = (C*)c.a + (long long)slot * c.bt.stride;
  C* ny = base + c.nyoff;
  double* part = c.partial + (long long)slot * c.bt.pzstride;
  C* lds = reinterpret_cast<C*>(rag_smem);
  // H = n2 / 2 <= 512 on every supported grid
#define XPFC_RZ(HH)                                                                               \
  case HH:                                                                                        \
    if constexpr (z_ok<T, HH, RAG_NTZ>())                                                         \
      zpass_body<T, HH, RAG_NTZ, Z_STEP, SAMPLE, PW>(base, ny, base, ny, c.nlines, c.zl,           \
                                                      c.bt.ia[slot].za, (const C*)c.tw,           \
                                                      (const C*)c.twp, part, lds, wb);            \
    break;
  switch (c.L) {
    XPFC_RZ(4) XPFC_RZ(8) XPFC_RZ(16) XPFC_RZ(32) XPFC_RZ(64) XPFC_RZ(128) XPFC_RZ(256) XPFC_RZ(512)
    default: break;
  }
#undef XPFC_RZ
}

namespace {
// dynamic LDS beyond the default limit needs the per-kernel attribute
template <class K>
bool allow_lds(K kernel, size_t bytes) {
  if (bytes <= 65536) return true;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)bytes) == hipSuccess;
}
template <typename T>
int rag_prepare_T(size_t lx, size_t ly, size_t lz) {
  const bool ok = allow_lds(k_xstep_r<T, true, true>, lx) && allow_lds(k_xstep_r<T, false, true>, lx) &&
                  allow_lds(k_xstep_r<T, true, false>, lx) && allow_lds(k_xstep_r<T, false, false>, lx) &&
                  allow_lds(k_strided_r<T, -1>, ly) && allow_lds(k_strided_r<T, 1>, ly) &&
                  allow_lds(k_zpass_r<T, true, true>, lz) && allow_lds(k_zpass_r<T, true, false>, lz) &&
                  allow_lds(k_zpass_r<T, false, true>, lz) && allow_lds(k_zpass_r<T, false, false>, lz);
  return ok ? 0 : launch_status(-2);
}
}  // namespace

int rag_prepare(int prec, size_t lx, size_t ly, size_t lz) {
  return prec == 1 ? rag_prepare_T<float>(lx, ly, lz) : rag_prepare_T<double>(lx, ly, lz);
}

template <typename T, bool LUT>
static int xr_T(const RClass* cls, int ncls, int nwg, size_t lds, bool sample, int sib, hipStream_t s) {
  if (nwg <= 0) return 0;
  if (sample)
    hipLaunchKernelGGL((k_xstep_r<T, true, LUT>), dim3(nwg), dim3(RAG_NT), lds, s, cls, ncls, sib);
  else
    hipLaunchKernelGGL((k_xstep_r<T, false, LUT>), dim3(nwg), dim3(RAG_NT), lds, s, cls, ncls, sib);
  return launch_status();
}

int launch_xstep_ragged(int prec, const RClass* cls, int ncls, int nwg, size_t lds, bool lut, bool sample,
                        int sib, hipStream_t s) {
  if (prec == 1)
    return lut ? xr_T<float, true>(cls, ncls, nwg, lds, sample, sib, s)
               : xr_T<float, false>(cls, ncls, nwg, lds, sample, sib, s);
  return lut ? xr_T<double, true>(cls, ncls, nwg, lds, sample, sib, s)
             : xr_T<double, false>(cls, ncls, nwg, lds, sample, sib, s);
}

template <typename T>
static int yr_T(const RClass* cls, int ncls, int nwg, size_t lds, int dir, int sib, hipStream_t s) {
  if (nwg <= 0) return 0;
  if (dir < 0)
    hipLaunchKernelGGL((k_strided_r<T, -1>), dim3(nwg), dim3(RAG_NT), lds, s, cls, ncls, sib);
  else
    hipLaunchKernelGGL((k_strided_r<T, 1>), dim3(nwg), dim3(RAG_NT), lds, s, cls, ncls, sib);
  return launch_status();
}

int launch_strided_ragged(int prec, const RClass* cls, int ncls, int nwg, size_t lds, int dir, int sib,
                          hipStream_t s) {
  return prec == 1 ? yr_T<float>(cls, ncls, nwg, lds, dir, sib, s) : yr_T<double>(cls, ncls, nwg, lds, dir, sib, s);
}

template <typename T, bool SAMPLE, bool PW>
static int zr_go(const RClass* cls, int ncls, int nwg, size_t lds, int sib, hipStream_t s) {
  hipLaunchKernelGGL((k_zpass_r<T, SAMPLE, PW>), dim3(nwg), dim3(RAG_NTZ), lds, s, cls, ncls, sib);
  return launch_status();
}

template <typename T>
static int zr_T(const RClass* cls, int ncls, int nwg, size_t lds, bool sample, bool pw, int sib, hipStream_t s) {
  if (nwg <= 0) return 0;
  if (sample) return pw ? zr_go<T, true, true>(cls, ncls, nwg, lds, sib, s) : zr_go<T, true, false>(cls, ncls, nwg, lds, sib, s);
  return pw ? zr_go<T, false, true>(cls, ncls, nwg, lds, sib, s) : zr_go<T, false, false>(cls, ncls, nwg, lds, sib, s);
}

int launch_zpass_ragged(int prec, const RClass* cls, int ncls, int nwg, size_t lds, bool sample, bool pw,
                        int sib, hipStream_t s) {
  return prec == 1 ? zr_T<float>(cls, ncls, nwg, lds, sample, pw, sib, s)
                   : zr_T<double>(cls, ncls, nwg, lds, sample, pw, sib, s);
}

template <typename T>
static int rsw_T(int L, const Segs& segs) {
  XPFC_RLSWITCH(L, {
    constexpr int TC = rag_tile(LL);
    if constexpr (tile_ok<T, LL, TC>()) {
      const Segs sg = fill_tiles(segs, TC);
      return sg.tiles0 + sg.tiles1;
    }
    return -3;
  })
  return -1;
}
int rag_strided_wgs(int prec, int L, const Segs& segs) {
  return prec == 1 ? rsw_T<float>(L, segs) : rsw_T<double>(L, segs);
}
size_t rag_strided_lds(int prec, int L) {
  if (L < 8) return 0;
  return ((size_t)L * (rag_tile(L) + PADT) + L) * (prec == 1 ? 8 : 16);
}

// (the half-line lengths k_zpass_r compiles: H = 4 .. 512)
template <typename T>
static long long rzw_T(int H, long long nlines, size_t* lds) {
  switch (H) {
#define XPFC_RZW(HH)                                                               \
  case HH:                                                                         \
    if constexpr (z_ok<T, HH, RAG_NTZ>()) {                                        \
      constexpr int G = RAG_NTZ / z_tpl<T, HH>();                                  \
      if (lds) *lds = (size_t)z_lds_elems<T, HH, RAG_NTZ>() * sizeof(cpx<T>);     \
      return (nlines + G - 1) / G;                                                 \
    }                                                                              \
    return -3;
    XPFC_RZW(4) XPFC_RZW(8) XPFC_RZW(16) XPFC_RZW(32) XPFC_RZW(64) XPFC_RZW(128) XPFC_RZW(256) XPFC_RZW(512)
#undef XPFC_RZW
    default: return -1;
  }
}
int rag_zpass_wgs(int prec, int H, long long nlines) {
  return (int)(prec == 1 ? rzw_T<float>(H, nlines, nullptr) : rzw_T<double>(H, nlines, nullptr));
}
size_t rag_zpass_lds(int prec, int H) {
  size_t b = 0;
  if ((prec == 1 ? rzw_T<float>(H, 1, &b) : rzw_T<double>(H, 1, &b)) < 0) return 0;
  return b;
}

}  // namespace xpfc
