// Line lengths of the step passes and their per-length launchers.
//
// Every length is compiled in one of the instantiation units xpfc_len_*.hip
// (explicit instantiations of LenOps, whose member definitions live in
// xpfc_lenops.hpp); xpfc_passes.hip dispatches on the runtime length and sees
// only these declarations, so the units compile in parallel.
#pragma once
#include "xpfc_internal.hpp"

namespace xpfc {

// Strided lines (n0, n1) and real-space half lines (H = n2 / 2):
// 2^k, 3 * 2^k and 5 * 2^k points.
#define XPFC_LENGTHS(X)                                                                                 \
  X(4) X(6) X(8) X(10) X(12) X(16) X(20) X(24) X(32) X(40) X(48) X(64) X(80) X(96) X(128) X(160)        \
  X(192) X(256) X(320) X(384) X(512) X(640) X(768) X(1024) X(1280) X(1536) X(2048)

// Launchers of one line length L (a strided pass over lines of L points, or
// the real-space pass over half lines of H = L points).  Return values as the
// launch_* entry points of xpfc_internal.hpp: -3 for a variant the length
// does not support.
template <typename T, int L>
struct LenOps {
  static int strided(const void* in, void* buf, const Segs& segs, int tc, int dir, const void* tw,
                     const Batch& bt, int ny, bool nt, hipStream_t s);
  static int blocks(int tc, const Segs& segs);
  static int xstep(void* work, void* psi, const Segs& segs, int tc, const void* tw, const KGeom& g,
                   const Batch& bt, int ny, double* partial, bool sample, hipStream_t s);
  static int zpass(int mode, const void* in, const void* in_ny, void* out, void* out_ny, long long nlines,
                   const ZLines& zl, int ntz, const Batch& bt, int ny, const void* tw, const void* twp,
                   double* partial, bool sample, bool pw, hipStream_t s);
  static int zblocks(int ntz, long long nlines);
};

}  // namespace xpfc
