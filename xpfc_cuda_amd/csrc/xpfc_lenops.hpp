// Per-length launchers of the step passes (definitions of LenOps, included
// only by the instantiation units xpfc_len_*.hip).
#pragma once
#include "xpfc_kernels.hpp"
#include "xpfc_lengths.hpp"

namespace xpfc {

#define XPFC_TCSWITCH(TCV, ...)                             \
  switch (TCV) {                                            \
    case 2: { constexpr int TT = 2; __VA_ARGS__; } break;   \
    case 4: { constexpr int TT = 4; __VA_ARGS__; } break;   \
    case 8: { constexpr int TT = 8; __VA_ARGS__; } break;   \
    case 16: { constexpr int TT = 16; __VA_ARGS__; } break; \
    default: return -3;                                     \
  }
#define XPFC_NTZSWITCH(V, ...)                                \
  switch (V) {                                                \
    case 64: { constexpr int NZ = 64; __VA_ARGS__; } break;   \
    case 128: { constexpr int NZ = 128; __VA_ARGS__; } break; \
    case 256: { constexpr int NZ = 256; __VA_ARGS__; } break; \
    case 512: { constexpr int NZ = 512; __VA_ARGS__; } break; \
    default: return -3;                                       \
  }

template <typename T, int LL, int TT, int DIR>
void strided_go(const void* in, void* buf, const Segs& sg, int nb, const void* tw, const Batch& bt, int ny,
                bool nt, hipStream_t s) {
  if (nt)
    hipLaunchKernelGGL((k_strided<T, LL, TT, DIR, true>), dim3(nb, ny), dim3(tile_threads<LL, TT>()), 0, s,
                       (const cpx<T>*)in, (cpx<T>*)buf, sg, (const cpx<T>*)tw, bt);
  else
    hipLaunchKernelGGL((k_strided<T, LL, TT, DIR, false>), dim3(nb, ny), dim3(tile_threads<LL, TT>()), 0, s,
                       (const cpx<T>*)in, (cpx<T>*)buf, sg, (const cpx<T>*)tw, bt);
}

template <typename T, int LL>
int LenOps<T, LL>::strided(const void* in, void* buf, const Segs& segs0, int tc, int dir, const void* tw,
                           const Batch& bt, int ny, bool nt, hipStream_t s) {
  XPFC_TCSWITCH(tc, {
    if constexpr (tile_ok<T, LL, TT>()) {
      const Segs sg = fill_tiles(segs0, TT);
      const int nb = sg.tiles0 + sg.tiles1;
      if (nb == 0 || ny == 0) return 0;
      if (dir < 0) strided_go<T, LL, TT, -1>(in, buf, sg, nb, tw, bt, ny, nt, s);
      else strided_go<T, LL, TT, +1>(in, buf, sg, nb, tw, bt, ny, nt, s);
    } else {
      return -3;
    }
  });
  return launch_status();
}

template <typename T, int LL>
int LenOps<T, LL>::blocks(int tc, const Segs& segs0) {
  XPFC_TCSWITCH(tc, {
    if constexpr (tile_ok<T, LL, TT>()) {
      const Segs sg = fill_tiles(segs0, TT);
      return sg.tiles0 + sg.tiles1;
    } else {
      return -3;
    }
  });
  return -1;
}

// tc: tile width; the square / cubic table path is taken when g.cubic is set.
// Returns the workgroup count (= number of X partials).
template <typename T, int LL>
int LenOps<T, LL>::xstep(void* work, void* psi, const Segs& segs0, int tc, const void* tw, const KGeom& g,
                         const Batch& bt, int ny, double* partial, bool sample, hipStream_t s) {
  int nb = 0;
  XPFC_TCSWITCH(tc, {
    if constexpr (tile_ok<T, LL, TT>()) {
      const Segs sg = fill_tiles(segs0, TT);
      nb = sg.tiles0 + sg.tiles1;
      if (ny == 0) return nb;
      const dim3 grid(nb, ny), block(tile_threads<LL, TT>());
      cpx<T>* w = (cpx<T>*)work;
      cpx<T>* p = (cpx<T>*)psi;
      const cpx<T>* t = (const cpx<T>*)tw;
      if (g.cubic) {
        if (sample)
          hipLaunchKernelGGL((k_xstep<T, LL, TT, true, true>), grid, block, 0, s, w, p, sg, t, g, bt, partial);
        else
          hipLaunchKernelGGL((k_xstep<T, LL, TT, false, true>), grid, block, 0, s, w, p, sg, t, g, bt, partial);
      } else {
        if (sample)
          hipLaunchKernelGGL((k_xstep<T, LL, TT, true, false>), grid, block, 0, s, w, p, sg, t, g, bt, partial);
        else
          hipLaunchKernelGGL((k_xstep<T, LL, TT, false, false>), grid, block, 0, s, w, p, sg, t, g, bt, partial);
      }
    } else {
      return -3;
    }
  });
  return launch_status(nb);
}

template <typename T, int H, int NZ, bool FRE, int MODE, bool PW = false>
void zlaunch(const void* in, const void* in_ny, void* out, void* out_ny, long long nlines, const ZLines& zl,
             const Batch& bt, int ny, const void* tw, const void* twp, double* partial, bool sample,
             hipStream_t s) {
  constexpr int G = NZ / z_tpl<T, H>();
  const dim3 grid((unsigned)((nlines + G - 1) / G), ny);
  if (sample)
    hipLaunchKernelGGL((k_zpass<T, H, NZ, FRE, MODE, true, PW>), grid, dim3(NZ), 0, s, in, (const cpx<T>*)in_ny,
                       out, (cpx<T>*)out_ny, nlines, zl, bt, (const cpx<T>*)tw, (const cpx<T>*)twp, partial);
  else
    hipLaunchKernelGGL((k_zpass<T, H, NZ, FRE, MODE, false, PW>), grid, dim3(NZ), 0, s, in,
                       (const cpx<T>*)in_ny, out, (cpx<T>*)out_ny, nlines, zl, bt, (const cpx<T>*)tw,
                       (const cpx<T>*)twp, partial);
}

// ntz: workgroup size (64..512); +ZFREE_REGS drops the occupancy hint.
// pw: some instance forms the FP32 powers (then the kernels that carry both
// forms run, without the register-allocation variant)
template <typename T, int LL>
int LenOps<T, LL>::zpass(int mode, const void* in, const void* in_ny, void* out, void* out_ny, long long nlines,
                         const ZLines& zl, int ntz, const Batch& bt, int ny, const void* tw, const void* twp,
                         double* partial, bool sample, bool pw, hipStream_t s) {
  if (ny == 0) return 0;
  const bool fre = (ntz & ZFREE_REGS) != 0;
  ntz &= ~ZFREE_REGS;
  if (mode == Z_STEP) {
    XPFC_NTZSWITCH(ntz, {
      if constexpr (z_ok<T, LL, NZ>()) {
        if (pw)
          zlaunch<T, LL, NZ, false, Z_STEP, true>(in, in_ny, out, out_ny, nlines, zl, bt, ny, tw, twp, partial,
                                                 sample, s);
        else if (fre)
          zlaunch<T, LL, NZ, true, Z_STEP>(in, in_ny, out, out_ny, nlines, zl, bt, ny, tw, twp, partial, sample,
                                           s);
        else
          zlaunch<T, LL, NZ, false, Z_STEP>(in, in_ny, out, out_ny, nlines, zl, bt, ny, tw, twp, partial, sample,
                                            s);
      } else {
        return -3;
      }
    });
  } else {
    // set-up and read-back modes: one workgroup shape (256 threads)
    constexpr int NZ = 256;
    if constexpr (z_ok<T, LL, NZ>()) {
      switch (mode) {
        case Z_R2C_FLUCT:
          zlaunch<T, LL, NZ, false, Z_R2C_FLUCT>(in, in_ny, out, out_ny, nlines, zl, bt, ny, tw, twp, partial,
                                                 false, s);
          break;
        case Z_R2C_NL:
          if (pw)
            zlaunch<T, LL, NZ, false, Z_R2C_NL, true>(in, in_ny, out, out_ny, nlines, zl, bt, ny, tw, twp,
                                                      partial, false, s);
          else
            zlaunch<T, LL, NZ, false, Z_R2C_NL>(in, in_ny, out, out_ny, nlines, zl, bt, ny, tw, twp, partial,
                                                false, s);
          break;
        case Z_C2R:
          if (pw)
            zlaunch<T, LL, NZ, false, Z_C2R, true>(in, in_ny, out, out_ny, nlines, zl, bt, ny, tw, twp, partial,
                                                   sample, s);
          else
            zlaunch<T, LL, NZ, false, Z_C2R>(in, in_ny, out, out_ny, nlines, zl, bt, ny, tw, twp, partial,
                                             sample, s);
          break;
        default: return -1;
      }
    } else {
      return -3;
    }
  }
  return launch_status();
}

template <typename T, int LL>
int LenOps<T, LL>::zblocks(int ntz, long long nlines) {
  ntz &= ~ZFREE_REGS;
  int G = 0;
  XPFC_NTZSWITCH(ntz, {
    if constexpr (z_ok<T, LL, NZ>()) G = NZ / z_tpl<T, LL>();
    else return -3;
  });
  return (int)((nlines + G - 1) / G);
}

#define XPFC_INSTANTIATE_LEN(L)         \
  template struct LenOps<float, L>;     \
  template struct LenOps<double, L>;

}  // namespace xpfc
