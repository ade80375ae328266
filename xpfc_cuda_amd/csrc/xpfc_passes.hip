// HIP kernels of the XPFC time step for gfx950 (MI355X): the dispatch of the
// step passes on the runtime line length (the per-length launchers are
// compiled in xpfc_len_*.hip), the fixed-order reductions, the C2 table, and
// the dispatch of the ragged launches that step instances of several grids in
// one launch per pass and line family (compiled in xpfc_rag_*.hip).  The pass
// kernels themselves are in xpfc_kernels.hpp and xpfc_ragged.hpp.
#include <climits>
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
int strided_T(const void* in, void* buf, const Segs& segs, int L, int tc, int dir, const void* tw,
              const Batch& bt, int ny, bool nt, hipStream_t s) {
  switch (L) {
#define XPFC_C(LEN) \
  case LEN: return LenOps<T, LEN>::strided(in, buf, segs, tc, dir, tw, bt, ny, nt, s);
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

int launch_strided(int prec, const void* in, void* buf, const Segs& segs, int L, int tc, int dir, const void* tw,
                   const Batch& bt, int ny, bool nt, hipStream_t s) {
  return prec == 1 ? strided_T<float>(in, buf, segs, L, tc, dir, tw, bt, ny, nt, s)
                   : strided_T<double>(in, buf, segs, L, tc, dir, tw, bt, ny, nt, s);
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

// C2 rows of a non-square 2-D box, one workgroup per kz: k^2 formed as the
// axis-0 pass forms it (k_y = 0), so each entry is the value that pass would
// evaluate for the mode
__global__ __launch_bounds__(256) void k_c2rows(double* __restrict__ tab, const C2Row* __restrict__ rows,
                                                double dk0, double dk2, Phys ph) {
  const int kz = blockIdx.x;
  const C2Row r = rows[kz];
  const double ky = 0.0;
  const double kzv = dk2 * (double)kz;
  const double kyz2 = ky * ky + kzv * kzv;
  for (int i = threadIdx.x; i < r.len; i += 256) {
    const double kx = dk0 * (double)i;
    const double k2 = kx * kx + kyz2;
    tab[r.off + i] = (k2 < ph.kcut2) ? c2hat(sqrt(k2), ph) : 0.0;
  }
}

int launch_c2rows(double* tab, const C2Row* rows, int nl, double dk0, double dk2, const Phys& ph, hipStream_t s) {
  if (nl <= 0) return 0;
  hipLaunchKernelGGL(k_c2rows, dim3(nl), dim3(256), 0, s, tab, rows, dk0, dk2, ph);
  return launch_status();
}

int launch_c2table(double* tab, int mcut, double dksq, const Phys& ph, hipStream_t s) {
  if (mcut <= 0) return 0;
  hipLaunchKernelGGL(k_c2table, dim3((mcut + 255) / 256), dim3(256), 0, s, tab, mcut, dksq, ph);
  return launch_status();
}

// ---------------------------------------------------------------------------
// ragged launches: several size classes in one launch per pass and line
// family (kernels in xpfc_ragged.hpp, one unit per family)
// ---------------------------------------------------------------------------
namespace {
// F(std::integral_constant<int, FAM>) for fam in {1, 3, 5}
template <class F>
int by_family(int fam, F&& f) {
  switch (fam) {
    case 1: return f(std::integral_constant<int, 1>{});
    case 3: return f(std::integral_constant<int, 3>{});
    case 5: return f(std::integral_constant<int, 5>{});
    default: return -3;
  }
}
// the family of a real-space half line: H = n2 / 2 of a 2^k, 3*2^k or 5*2^k side
int half_family(int H) { return line_family(H); }
}  // namespace

int rag_prepare(int prec, int fam) {
  return by_family(fam, [&](auto f) {
    constexpr int F = decltype(f)::value;
    return prec == 1 ? RagOps<float, F>::prepare() : RagOps<double, F>::prepare();
  });
}

int launch_xstep_ragged(int prec, int fam, const RClass* cls, int ncls, int nwg, size_t lds, bool lut,
                        bool sample, int sib, hipStream_t s) {
  return by_family(fam, [&](auto f) {
    constexpr int F = decltype(f)::value;
    return prec == 1 ? RagOps<float, F>::xstep(cls, ncls, nwg, lds, lut, sample, sib, s)
                     : RagOps<double, F>::xstep(cls, ncls, nwg, lds, lut, sample, sib, s);
  });
}

int launch_strided_ragged(int prec, int fam, const RClass* cls, int ncls, int nwg, size_t lds, int dir,
                          int sib, hipStream_t s) {
  return by_family(fam, [&](auto f) {
    constexpr int F = decltype(f)::value;
    return prec == 1 ? RagOps<float, F>::strided(cls, ncls, nwg, lds, dir, sib, s)
                     : RagOps<double, F>::strided(cls, ncls, nwg, lds, dir, sib, s);
  });
}

int launch_zpass_ragged(int prec, int fam, const RClass* cls, int ncls, int nwg, size_t lds, bool sample,
                        bool pw, int sib, hipStream_t s) {
  return by_family(fam, [&](auto f) {
    constexpr int F = decltype(f)::value;
    return prec == 1 ? RagOps<float, F>::zpass(cls, ncls, nwg, lds, sample, pw, sib, s)
                     : RagOps<double, F>::zpass(cls, ncls, nwg, lds, sample, pw, sib, s);
  });
}

int rag_strided_wgs(int prec, int L, const Segs& segs) {
  if (L < 8) return -1;
  return by_family(line_family(L), [&](auto f) {
    constexpr int F = decltype(f)::value;
    return prec == 1 ? RagOps<float, F>::strided_wgs(L, segs) : RagOps<double, F>::strided_wgs(L, segs);
  });
}
size_t rag_strided_lds(int prec, int L) {
  if (L < 8) return 0;
  return ((size_t)L * (rag_tile(L) + PADT) + L) * (prec == 1 ? 8 : 16);
}

namespace {
long long rzw(int prec, int H, long long nlines, size_t* lds) {
  if (H < 4) return -1;
  return by_family(half_family(H), [&](auto f) -> int {
    constexpr int F = decltype(f)::value;
    const long long n = prec == 1 ? RagOps<float, F>::zpass_wgs(H, nlines, lds)
                                  : RagOps<double, F>::zpass_wgs(H, nlines, lds);
    return n > INT_MAX ? -1 : (int)n;
  });
}
}  // namespace
int rag_zpass_wgs(int prec, int H, long long nlines) { return (int)rzw(prec, H, nlines, nullptr); }
size_t rag_zpass_lds(int prec, int H) {
  size_t b = 0;
  if (rzw(prec, H, 1, &b) < 0) return 0;
  return b;
}

}  // namespace xpfc
