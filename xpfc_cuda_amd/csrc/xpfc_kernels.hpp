// Step-pass kernels of the XPFC time step for gfx950 (MI355X): the device
// bodies and __global__ templates shared by the per-length instantiation units
// (xpfc_len_*.hip) and the ragged launches (xpfc_passes.hip).
//
// Spectrum layout (packed half spectrum): for a real field of n0 x n1 x n2
// points (n2 contiguous; 2-D fields use n1 = 1) the spectrum is
//   main[n0][n1][H]  (kz = 0 .. H-1, H = n2/2)  followed by  ny[n0][n1] (kz = H).
// The k = 0 mode is always 0: the mean density is carried separately in FP64
// and added back in real space.
//
// The state between steps, in the `psi` buffer, takes one of two forms (per
// context, KGeom::xreal): psi_k itself, or psi_k transformed back along axis 0
// (the unnormalised inverse transform over kx; spectral along the other axes).
// `work` carries the nonlinear term N_k transformed along every axis but
// axis 0.  One time step = four passes over HBM (two in 2-D):
//   k_xstep   : lines along axis 0: forward FFT of the nonlinear term (and of
//               the state in the second form: psi_k = FFT_0(state) / n0), the
//               semi-implicit update of psi_k, inverse FFT
//                                                             (euler_step.cu:50-59)
//               k-space form: psi_k' and its inverse transform (into work)
//               are written, 4 spectra of traffic; axis-0 form: only the
//               inverse transform, which is the new state, 3 spectra
//   k_strided : lines along axis 1, inverse (axis-0 form: state -> work)
//                                                                 (3-D only)
//   k_zpass   : lines along axis 2: C2R, 1/N, N(psi) = -a psi^2 + b psi^3,
//               R2C (2-D axis-0 form: state -> work)      (power_functions.cu:34-53,
//                                                                  scaling_step.cu:48-49,
//                                                                  copy_matrix.cu:44-45)
//   k_strided : lines along axis 1, forward                      (3-D only)
// The axis-0 form saves one spectrum of traffic per step at the price of one
// more transform of the state; it pays where the spectra stream from HBM and
// costs where they are cache-resident (small grids), hence the choice.
// The sampled diagnostics (FE_calc.cu:64, sse_calculation.cu:133-140,
// Avg_calc.cu:3-18) are FP64 partial sums fused into k_xstep and k_zpass.
//
// Line lengths: 2^k, 3*2^k and 5*2^k points (the FFT stages of fft_core.hpp
// with radices up to 8, 12 and 20).
//
// Variants: the strided kernels take the tile width TC (columns per
// workgroup) and the real-space kernel its workgroup size NTZ as template
// parameters; the order in which workgroups are dealt to the XCDs and the
// streaming cache hints of the axis-1 passes are further options.  They change
// only how lines are grouped, ordered and cached, never the arithmetic of a
// line, so every variant gives bitwise-identical fields; the autotuner picks
// the fastest per size.
#pragma once
#include <type_traits>

#include "xpfc_device.hpp"

namespace xpfc {

constexpr int PADT = 1;  // LDS padding (complex elements) per tile row

template <int L, int TC> constexpr int tile_threads() { return TC * (L / elems_per_thread<L>()); }
// LDS: the tile plus the L-entry twiddle table
template <typename T, int L, int TC> constexpr bool tile_fits() {
  return tile_threads<L, TC>() <= 1024 &&
         (long long)L * (TC + PADT + 1) * (long long)sizeof(cpx<T>) <= 152 * 1024;
}
// 2-column tiles only where 4 columns do not fit (FP64 lines of 2048 points)
template <typename T, int L, int TC> constexpr bool tile_ok() {
  return TC == 2 ? (!tile_fits<T, L, 4>() && tile_fits<T, L, 2>()) : tile_fits<T, L, TC>();
}

// XCD-aware block order: the dispatcher deals consecutive workgroups to the 8
// XCDs round-robin.  lgc < 0, or fewer tiles than one round: each XCD receives
// one contiguous run of tiles.  Otherwise each XCD receives runs of 2^lgc
// neighbouring tiles (column tiles that share 128-byte lines share one L2),
// the runs dealt to the XCDs in turn, so that the eight XCDs work on nearby
// columns at any time; tiles past the last whole round keep their order.
__device__ __forceinline__ int xcd_tile(int b, int n, int lgc) {
  constexpr int X = 8;
  const int xcd = b % X, slot = b / X;
  if (lgc >= 0 && n >= (X << lgc)) {
    const int full = (n >> (lgc + 3)) << (lgc + 3);
    if (b >= full) return b;
    return ((((slot >> lgc) * X) + xcd) << lgc) + (slot & ((1 << lgc) - 1));
  }
  const int per = n / X, rem = n % X;
  return xcd < rem ? xcd * (per + 1) + slot : rem * (per + 1) + (xcd - rem) * per + slot;
}

// The tile counts of a strided pass with TC-column tiles.
inline Segs fill_tiles(const Segs& in, int TC) {
  Segs s = in;
  int tiles[2];
  for (int i = 0; i < 2; ++i) {
    s.s[i].ntc = (s.s[i].ncols + TC - 1) / TC;
    tiles[i] = (s.s[i].ncols <= 0 || s.s[i].nbatch <= 0) ? 0 : s.s[i].ntc * s.s[i].nbatch;
  }
  s.tiles0 = tiles[0];
  s.tiles1 = tiles[1];
  return s;
}

// Where a tile's column lies: its segment's row stride and kind are copied
// out (no pointer into the kernel arguments, which would force a private copy
// of them once the location is carried across loop iterations).
struct TileLoc {
  int col;
  int rs;
  int kind;
  bool act;
  long long off;
};

__device__ __forceinline__ TileLoc locate(const Segs& segs, int tile, int ci, int TC) {
  TileLoc r;
  const bool second = tile >= segs.tiles0;
  if (second) tile -= segs.tiles0;
  const Seg& sg = second ? segs.s[1] : segs.s[0];
  const int b = tile / sg.ntc, ct = tile - b * sg.ntc;
  r.col = ct * TC + ci;
  r.act = r.col < sg.ncols;
  r.off = sg.off + (long long)b * sg.bs + (long long)r.col * sg.cs;
  r.rs = sg.rs;
  r.kind = sg.kind;
  return r;
}

// ---------------------------------------------------------------------------
// strided pass: forward (DIR=-1) or inverse (DIR=+1) transform of every line
// ---------------------------------------------------------------------------
// NT: loads and stores carry the streaming hint.  For fields larger than the
// last-level cache the passes that follow the axis-1 passes then run faster
// (measured at 512^3: the axis-0 pass after the forward axis-1 pass -6 %, the
// real-space pass after the inverse one -6 %, the forward axis-1 pass itself
// +7 %; the step -2.3 % in FP32, -4 % in FP64); for fields that fit the cache
// the hint costs up to 15 %.
// One column tile of a strided pass (buf already at the instance's spectra).
// `in` and `out` are the same buffer (in place) or two buffers of one layout:
// every line is read before the transform and written after it.
template <typename T, int L, int TC, int DIR, bool NT>
__device__ __forceinline__ void strided_body(const cpx<T>* in, cpx<T>* out, const Segs& segs, int tile,
                                             const cpx<T>* __restrict__ tw, cpx<T>* lds, cpx<T>* twl) {
  using C = cpx<T>;
  constexpr int E = elems_per_thread<L>();
  const int ci = threadIdx.x % TC, t = threadIdx.x / TC;
  const TileLoc loc = locate(segs, tile, ci, TC);
  const int rs = loc.rs;
  // row t + m * L/E of the column: the thread's offset plus a uniform step
  // (t * rs and (L/E) * rs stay below 2^31 on every supported grid; m times
  // the step may not, at 2048 points along axis 0 and large planes)
  const int r0 = t * rs, rstep = (L / E) * rs;
  const C* src = in + loc.off + r0;
  C* p = out + loc.off + r0;
  C v[E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const C* q = src + (long long)m * rstep;
    v[m] = loc.act ? (NT ? load_nt(q) : *q) : mk<C>(0, 0);
  }
  stage_table<tile_threads<L, TC>()>(twl, tw, L);  // read after the first exchange's barrier
  fft_line<T, L, E, DIR>(v, t, lds, TileIdx<TC, PADT>{ci}, twl);
  if (loc.act) {
#pragma unroll
    for (int m = 0; m < E; ++m) {
      C* q = p + (long long)m * rstep;
      if constexpr (NT) store_nt(q, v[m]);
      else *q = v[m];
    }
  }
}

template <typename T, int L, int TC, int DIR, bool NT>
__global__ __launch_bounds__((tile_threads<L, TC>()))
void k_strided(const cpx<T>* in, cpx<T>* out, Segs segs, const cpx<T>* __restrict__ tw, Batch bt) {
  using C = cpx<T>;
  __shared__ C lds[L * (TC + PADT)];
  __shared__ C twl[L];
  const int slot = bt.slots[blockIdx.y];
  if (bt.budget && bt.sib >= bt.budget[slot]) return;  // the whole workgroup
  const long long so = (long long)slot * bt.stride;
  strided_body<T, L, TC, DIR, NT>(in + so, out + so, segs, xcd_tile(blockIdx.x, gridDim.x, segs.lgc), tw, lds,
                                  twl);
}

// ---------------------------------------------------------------------------
// fused axis-0 pass: FFT(N) and FFT(state) along axis 0, semi-implicit
// update, inverse FFT into the state
// ---------------------------------------------------------------------------
// LUT: square / cubic grid, k^2 = dksq * m from the integer m = i^2 + j^2 + l^2
// and C2 from the instance's table over m (no FP64 exp in the kernel, which
// keeps the register footprint of the update small); otherwise k^2 and C2 are
// evaluated from the per-axis wavenumbers.
//
// Measured (512^3 FP32, rocprofv3 counters): this pass keeps fewer reads in
// flight than the one-stream passes and sees almost no DRAM credit stalls, so
// it is bound by how many requests it issues, not by HBM.  Raising occupancy
// by capping registers (80 or 64 VGPRs) spills and runs 20-150 % slower;
// walking 2 or 4 tiles per workgroup with the next tile's loads in flight
// needs 167-178 VGPRs; fetching the state or C2 after the forward transform
// is 4 % slower.  The variant below (one tile, both streams and C2 loaded up
// front) is the fastest measured.  (Measured with the k-space state, written
// back beside the working buffer; the state is now held along axis 0 in real
// space, which drops that write.)

// One tile's loads: the nonlinear term and the state along the column.
template <typename T, int L, int E>
struct XTile {
  cpx<T> v[E], ps[E];
  TileLoc loc;
};

template <typename T, int L, int TC>
__device__ __forceinline__ void xload(XTile<T, L, elems_per_thread<L>()>& x, const Segs& segs, int tile,
                                      int ntiles, int ci, int t, const cpx<T>* __restrict__ work,
                                      const cpx<T>* __restrict__ psi) {
  using C = cpx<T>;
  constexpr int E = elems_per_thread<L>();
  x.loc = locate(segs, tile < ntiles ? tile : 0, ci, TC);
  x.loc.act = x.loc.act && tile < ntiles;
  const int r0 = t * x.loc.rs, rstep = (L / E) * x.loc.rs;
  const C* pw = work + x.loc.off + r0;
  const C* pp = psi + x.loc.off + r0;
#pragma unroll
  for (int m = 0; m < E; ++m) x.v[m] = x.loc.act ? pw[(long long)m * rstep] : mk<C>(0, 0);
#pragma unroll
  for (int m = 0; m < E; ++m) x.ps[m] = x.loc.act ? pp[(long long)m * rstep] : mk<C>(0, 0);
}

template <typename T, int L, int TC, bool SAMPLE, bool LUT>
__device__ __forceinline__ void xtile(XTile<T, L, elems_per_thread<L>()>& x, int ci, int t,
                                      cpx<T>* __restrict__ work, cpx<T>* __restrict__ psi,
                                      cpx<T>* lds, const cpx<T>* twl, const KGeom& g, const Phys& ph,
                                      double (&acc)[XPART]) {
  using C = cpx<T>;
  constexpr int E = elems_per_thread<L>();
  const TileLoc& loc = x.loc;
  const int r0 = t * loc.rs, rstep = (L / E) * loc.rs;
  const bool act = loc.act;
  C* pp = psi + loc.off + r0;

  // wavenumber of this column
  int kb, kz;
  if (loc.kind == 0) { kb = loc.col / g.H; kz = loc.col - kb * g.H; }
  else { kb = loc.col; kz = g.H; }
  const int jb = signed_index(kb, g.n1);
  const int myz = jb * jb + kz * kz;  // LUT: integer part of k^2 from axes 1, 2
  // the line's C2 values are gathered with the input streams (gathering them
  // after the forward transform was measured slower, and for 12-point threads
  // it raised the register count).  Boxes that are not square / cubic evaluate
  // them here too, while the loads are in flight: inside the update loop the
  // FP64 exp temporaries met the transformed lines and took the 512-point FP32
  // pass from 117 to 136 VGPRs, one workgroup per CU instead of two (512^3 with
  // one side scaled: axis-0 pass 0.66 ms evaluated in the loop, 0.40 ms here).
  // (A per-m table of the update's FP32 coefficients for every mode, which
  // would save the FP64 chain, made the pass 22 % slower at 512^3: the C2
  // table is read only below the cutoff, 9 % of the modes, the full table for
  // every mode.)
  const double ky = g.dk1 * (double)jb;
  const double kzv = g.dk2 * (double)kz;
  const double kyz2 = ky * ky + kzv * kzv;
  double c2t[E];
  if constexpr (LUT) {
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int ia = signed_index(t + m * (L / E), g.n0);
      c2t[m] = c2_lookup(__mul24(ia, ia) + myz, ph);  // |ia| <= n0 / 2: 24-bit products
    }
  } else if (ph.c2rows) {  // 2-D box that is not square: the instance's rows
    C2Row r{0, 0};
    if (kz < ph.c2nl) r = ph.c2rows[kz];
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int ia = signed_index(t + m * (L / E), g.n0);
      const int ai = ia < 0 ? -ia : ia;
      c2t[m] = ai < r.len ? ph.c2tab[r.off + ai] : 0.0;
    }
  } else {
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int ia = signed_index(t + m * (L / E), g.n0);
      const double kx = g.dk0 * (double)ia;
      const double k2 = kx * kx + kyz2;
      c2t[m] = (k2 < ph.kcut2) ? c2hat(sqrt(k2), ph) : 0.0;
    }
  }
  C (&v)[E] = x.v;
  // the nonlinear term's spectrum, and in the axis-0 form the state's
  // (psi_k = FFT_0(state) / n0, the k = 0 mode exactly 0: the mean is carried
  // apart).  Transforming the two together through two LDS tiles (one barrier
  // pair per stage for both) was measured no faster in FP32 and 12 % slower in
  // FP64 at 512^3.
  const bool xreal = g.xreal != 0;
  fft_line<T, L, E, -1>(v, t, lds, TileIdx<TC, PADT>{ci}, twl);
  if (xreal) fft_line<T, L, E, -1>(x.ps, t, lds, TileIdx<TC, PADT>{ci}, twl);
  const T s0 = (T)g.inv_n0;
  const bool kzero = xreal && loc.kind == 0 && loc.col == 0;

  const double wgt = act ? ((loc.kind == 0 && kz == 0) || loc.kind == 1 ? 1.0 : 2.0) : 0.0;

  // psi_k' = (psi_k - dt M k^2 N_k) / den  (euler_step.cu:50-59).  The
  // denominator and its reciprocal are FP64; the numerator is field arithmetic
  // in the working precision.  Each element's FP64 chain is fenced off with a
  // scheduling barrier so the unrolled iterations do not interleave and the
  // register footprint stays that of one element.
#pragma unroll
  for (int m = 0; m < E; ++m) {
    __builtin_amdgcn_sched_barrier(0);
    const int ka = t + m * (L / E);
    const C old = !xreal ? x.ps[m] : (kzero && ka == 0) ? mk<C>(0, 0) : mk<C>(x.ps[m].x * s0, x.ps[m].y * s0);
    const int ia = signed_index(ka, g.n0);
    double k2, c2;
    if constexpr (LUT) {
      k2 = g.dksq * (double)(__mul24(ia, ia) + myz);
      c2 = c2t[m];
    } else {
      const double kx = g.dk0 * (double)ia;
      k2 = kx * kx + kyz2;
      c2 = c2t[m];
    }
    double a2, inv;
    update_coefs(k2, c2, ph, a2, inv);
    const C nh = v[m];
    const T ti = (T)inv, ta = (T)a2;
    const C nv = mk<C>(ti * (old.x - ta * nh.x), ti * (old.y - ta * nh.y));
    if constexpr (SAMPLE) {
      const double nx = (double)nv.x, ny = (double)nv.y;
      acc[0] += wgt * c2 * (nx * nx + ny * ny);
      // sse_calculation.cu:133-140 (eps absent from the residual, as there)
      const double rx = (nx - (double)old.x) * ph.inv_dt
                        + ph.Mres * k2 * ((double)old.x + (double)nh.x - c2 * (double)old.x);
      const double ry = (ny - (double)old.y) * ph.inv_dt
                        + ph.Mres * k2 * ((double)old.y + (double)nh.y - c2 * (double)old.y);
      acc[1] += wgt * (rx * rx + ry * ry);
    }
    v[m] = nv;
  }
  __builtin_amdgcn_sched_barrier(0);
  if (!xreal && act) {
#pragma unroll
    for (int m = 0; m < E; ++m) pp[(long long)m * rstep] = v[m];
  }

  // back along axis 0 (unnormalised): the new state in the axis-0 form, the
  // working buffer's input of the next passes in the k-space form
  fft_line<T, L, E, +1>(v, t, lds, TileIdx<TC, PADT>{ci}, twl);
  C* out = xreal ? pp : work + loc.off + r0;
  if (act) {
#pragma unroll
    for (int m = 0; m < E; ++m) out[(long long)m * rstep] = v[m];
  }
}

// One column tile of the fused axis-0 pass (work, psi, partial already at the
// instance's regions; `wb`: the workgroup's partial-sum slot).
template <typename T, int L, int TC, bool SAMPLE, bool LUT>
__device__ __forceinline__ void xstep_body(cpx<T>* __restrict__ work, cpx<T>* __restrict__ psi,
                                           const Segs& segs, int tile, const cpx<T>* __restrict__ tw,
                                           const KGeom& g, const Phys& ph, cpx<T>* lds, cpx<T>* twl,
                                           double* __restrict__ partial, int wb) {
  constexpr int E = elems_per_thread<L>();
  constexpr int NT = tile_threads<L, TC>();
  const int ci = threadIdx.x % TC, t = threadIdx.x / TC;
  const int ntiles = segs.tiles0 + segs.tiles1;
  double acc[XPART] = {0.0, 0.0};
  // both input streams are read up front (their latency hides behind the
  // forward transform), the two output streams written in bursts after it
  XTile<T, L, E> x;
  xload<T, L, TC>(x, segs, tile, ntiles, ci, t, work, psi);
  stage_table<NT>(twl, tw, L);  // read after the first exchange's barrier
  xtile<T, L, TC, SAMPLE, LUT>(x, ci, t, work, psi, lds, twl, g, ph, acc);
  if constexpr (SAMPLE) block_sum<NT, XPART>(acc, partial + (long long)wb * XPART);
}

template <typename T, int L, int TC, bool SAMPLE, bool LUT>
__global__ __launch_bounds__((tile_threads<L, TC>()))
void k_xstep(cpx<T>* __restrict__ work, cpx<T>* __restrict__ psi, Segs segs,
             const cpx<T>* __restrict__ tw, KGeom g, Batch bt, double* __restrict__ partial) {
  using C = cpx<T>;
  __shared__ C lds[L * (TC + PADT)];
  __shared__ C twl[L];
  const int slot = bt.slots[blockIdx.y];
  if (bt.budget && bt.sib >= bt.budget[slot]) return;
  work += (long long)slot * bt.stride;
  psi += (long long)slot * bt.stride;
  partial += (long long)slot * bt.pstride;
  xstep_body<T, L, TC, SAMPLE, LUT>(work, psi, segs, xcd_tile(blockIdx.x, gridDim.x, segs.lgc), tw, g,
                                    bt.ia[slot].ph, lds, twl, partial, blockIdx.x);
}

// ---------------------------------------------------------------------------
// real-space pass along the contiguous axis (H-point complex FFT of the
// even/odd packed line, with the half-spectrum pre/post twiddle)
// ---------------------------------------------------------------------------
template <typename T, int H> constexpr int z_tpl() { return H / z_elems<T, H>(); }  // threads per line
// (a line's lanes must lie in one wavefront: the twiddle pairs are exchanged
// with __shfl; every supported H = n2/2 satisfies it)
// LDS line of H rows plus the X[H] slot.  A Stockham stage writes rows E*j + q
// from lanes j.  For power-of-two E the rows are XOR-swizzled within each
// block of E, which spreads those lanes over distinct banks, and the odd line
// length puts neighbouring lines on the other half of the banks; for other E
// one padding slot follows every block of E rows (lane stride E + 1).
template <int H, int E> constexpr int z_len() { return (E & (E - 1)) == 0 ? H + 1 : H + H / E + 1; }
template <typename T, int H> constexpr int z_line() { return z_len<H, z_elems<T, H>()>(); }
template <int H, int E> struct ZIdx {
  int g;
  __device__ __forceinline__ int operator()(int r) const {
    if constexpr ((E & (E - 1)) == 0) {
      const int s = (r >= H) ? r : (r ^ ((r / E) & (E - 1)));
      return g * z_len<H, E>() + s;
    } else {
      return g * z_len<H, E>() + r + r / E;
    }
  }
};
template <typename T, int H, int NTZ> constexpr bool z_ok() {
  return z_tpl<T, H>() <= 64 && NTZ >= z_tpl<T, H>() && NTZ % z_tpl<T, H>() == 0 &&
         ((long long)(NTZ / z_tpl<T, H>()) * z_line<T, H>() + 2 * H) * (long long)sizeof(cpx<T>) <= 152 * 1024;
}

// Free-energy sums (psi, psi^2, psi^3, psi^4) of two real-space values, and
// the nonlinear term N(psi) = -a psi^2 + b psi^3.  F32: the powers are those
// of the reference's FP32 powf (power_functions.cu:47), FP32 products of
// float(psi); otherwise exact products (FP64 sums, working-precision N).
template <bool F32>
__device__ __forceinline__ void fe_sums(double (&acc)[ZPART], double p0, double p1) {
  acc[0] += p0 + p1;
  if constexpr (F32) {
    const float f0 = (float)p0, f1 = (float)p1;
    const float s0 = f0 * f0, s1 = f1 * f1;
    const float c0 = s0 * f0, c1 = s1 * f1;
    acc[1] += (double)s0 + (double)s1;
    acc[2] += (double)c0 + (double)c1;
    acc[3] += (double)(c0 * f0) + (double)(c1 * f1);
  } else {
    acc[1] += p0 * p0 + p1 * p1;
    acc[2] += p0 * p0 * p0 + p1 * p1 * p1;
    acc[3] += (p0 * p0) * (p0 * p0) + (p1 * p1) * (p1 * p1);
  }
}

template <bool F32, typename T>
__device__ __forceinline__ T nonlin(double p, T a, T b, const ZArgs& za) {
  if constexpr (F32) {
    const float f = (float)p, sq = f * f, cu = sq * f;
    return (T)(za.b * (double)cu - za.a * (double)sq);
  } else {
    const T q = (T)p;
    return q * q * (b * q - a);
  }
}

// Runs body(std::true_type) for instances with the FP32 powers, else
// body(std::false_type); PW = false compiles the default form only (a kernel
// with both forms holds more registers, so it is launched only for contexts
// that have an instance with the switch set).
template <bool PW, class F>
__device__ __forceinline__ void with_powers(const ZArgs& za, F&& body) {
  if constexpr (PW) {
    if (za.powf32) body(std::true_type{});
    else body(std::false_type{});
  } else {
    body(std::false_type{});
  }
}

// The C2R pre-twiddle and the R2C post-twiddle pair bin k with bin H-k; the
// mirror is exchanged with __shfl among the TPL lanes that hold the line (a
// power of two <= 64, so a line never straddles two wavefronts).
// FREE: no occupancy hint (the compiler may use more registers instead of
// spilling); a register-allocation variant only, the arithmetic is identical.
// LDS of the real-space pass: G lines of z_line<T, H> plus the two H-entry tables.
template <typename T, int H, int NTZ> constexpr int z_lds_elems() {
  return (NTZ / z_tpl<T, H>()) * z_line<T, H>() + 2 * H;
}

// The lines of workgroup `wb` of one instance (pointers already at the
// instance's regions); `lds` holds the line buffer and, after it, the tables.
template <typename T, int H, int NTZ, int MODE, bool SAMPLE, bool PW>
__device__ __forceinline__ void zpass_body(const void* __restrict__ in, const cpx<T>* __restrict__ in_ny,
                                           void* __restrict__ out, cpx<T>* __restrict__ out_ny,
                                           long long nlines, const ZLines& zl, const ZArgs& za,
                                           const cpx<T>* __restrict__ tw, const cpx<T>* __restrict__ twp,
                                           double* __restrict__ partial, cpx<T>* lds, int wb) {
  using C = cpx<T>;
  constexpr int E = z_elems<T, H>();
  constexpr int TPL = z_tpl<T, H>();
  static_assert(TPL <= 64 && (TPL & (TPL - 1)) == 0, "a line must lie in one wavefront");
  constexpr int G = NTZ / TPL;
  C* twl = lds + G * z_line<T, H>();  // the H-point FFT twiddles
  C* twpl = twl + H;              // and the half-spectrum twiddles
  stage_table<NTZ>(twl, tw, H);
  stage_table<NTZ>(twpl, twp, H);
  __syncthreads();
  const int tid = threadIdx.x;
  const int g = tid / TPL, t = tid - g * TPL;
  const long long line = (long long)wb * G + g;
  const bool act = line < nlines;
  // line = la * n1 + lb (lines per instance < 2^31)
  const long long la = (long long)((unsigned)line / (unsigned)zl.n1), lb = line - la * zl.n1;
  const long long off_in = la * zl.xs_in + lb * H, off_out = la * zl.xs_out + lb * H;
  const ZIdx<H, E> idx{g};
  double acc[ZPART] = {0.0, 0.0, 0.0, 0.0};
  C v[E];

  if constexpr (MODE == Z_STEP || MODE == Z_C2R) {
    const C* sp = reinterpret_cast<const C*>(in) + off_in;
    const T s = (T)za.scale;
    // The pre-twiddle pairs X[k] with X[H-k].  For t >= 1, X[H-k] (k = t + m*TPL)
    // is element E-1-m of lane TPL-t of the same line; for t = 0 it is this
    // lane's element E-m (and the Nyquist bin for m = 0).  The line is read
    // once and the mirror exchanged across lanes.
    C xo[E];
#pragma unroll
    for (int m = 0; m < E; ++m) xo[m] = act ? sp[t + m * TPL] : mk<C>(0, 0);
    const C xnyq = (act && t == 0) ? in_ny[line] : mk<C>(0, 0);
    const int lane = threadIdx.x & 63;
    const int src = lane - t + ((TPL - t) & (TPL - 1));
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int k = t + m * TPL;
      C xk = xo[m];
      C xm;
      xm.x = __shfl(xo[E - 1 - m].x, src, 64);
      xm.y = __shfl(xo[E - 1 - m].y, src, 64);
      if (t == 0) xm = (m == 0) ? xnyq : xo[m == 0 ? 0 : E - m];
      if (k == 0) { xk.y = 0; xm.y = 0; }  // DC and Nyquist bins are real (C2R)
      const C e = mk<C>(xk.x + xm.x, xk.y - xm.y);
      const C d = mk<C>(xk.x - xm.x, xk.y + xm.y);
      const C o = cmul(d, cconj(twpl[k]));
      v[m] = mk<C>((e.x - o.y) * s, (e.y + o.x) * s);
    }
    fft_line<T, H, E, +1>(v, t, lds, idx, twl);
    if constexpr (MODE == Z_C2R) {
      C* rp = reinterpret_cast<C*>(out) + off_out;  // real pairs
      with_powers<PW>(za, [&](auto f32) {
#pragma unroll
        for (int m = 0; m < E; ++m) {
          const int k = t + m * TPL;
          const double p0 = za.mean + (double)v[m].x, p1 = za.mean + (double)v[m].y;
          if constexpr (SAMPLE) {
            if (act) fe_sums<decltype(f32)::value>(acc, p0, p1);
          }
          if (act) rp[k] = mk<C>((T)p0, (T)p1);
        }
      });
      if constexpr (SAMPLE) block_sum<NTZ, ZPART>(acc, partial + (long long)wb * ZPART);
      return;
    } else {
      const T a = (T)za.a, b = (T)za.b;
      with_powers<PW>(za, [&](auto f32) {
        constexpr bool F = decltype(f32)::value;
#pragma unroll
        for (int m = 0; m < E; ++m) {
          const double p0 = za.mean + (double)v[m].x, p1 = za.mean + (double)v[m].y;
          if constexpr (SAMPLE) {
            if (act) fe_sums<F>(acc, p0, p1);
          }
          v[m] = mk<C>(nonlin<F, T>(p0, a, b, za), nonlin<F, T>(p1, a, b, za));
        }
      });
    }
  } else {
    // input: the fluctuation psi - mean as real pairs (x[2k], x[2k+1])
    const C* rp = reinterpret_cast<const C*>(in) + off_in;
    const T a = (T)za.a, b = (T)za.b;
    with_powers<PW>(za, [&](auto f32) {
      constexpr bool F = decltype(f32)::value;
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int k = t + m * TPL;
        C x = act ? rp[k] : mk<C>(0, 0);
        if constexpr (MODE == Z_R2C_NL)
          x = mk<C>(nonlin<F, T>(za.mean + (double)x.x, a, b, za), nonlin<F, T>(za.mean + (double)x.y, a, b, za));
        v[m] = act ? x : mk<C>(0, 0);
      }
    });
  }

  // forward H-point transform and the real-to-half-spectrum post twiddle,
  // which pairs Z[k] with Z[H-k]: the mirror bin is exchanged across the
  // line's lanes as in the pre-twiddle above (element E-1-m of lane TPL-t;
  // for t = 0 this lane's element E-m, or Z[0] itself)
  fft_line<T, H, E, -1>(v, t, lds, idx, twl);
  C* op = reinterpret_cast<C*>(out) + off_out;
  const T half = (T)0.5;
  const int mlane = threadIdx.x & 63;
  const int msrc = mlane - t + ((TPL - t) & (TPL - 1));
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int k = t + m * TPL;
    const C zk = v[m];
    C zm;
    zm.x = __shfl(v[E - 1 - m].x, msrc, 64);
    zm.y = __shfl(v[E - 1 - m].y, msrc, 64);
    if (t == 0) zm = v[m == 0 ? 0 : E - m];
    const C e = mk<C>(half * (zk.x + zm.x), half * (zk.y - zm.y));
    const C o = mk<C>(half * (zk.y + zm.y), -half * (zk.x - zm.x));
    if (act) {
      if (k == 0) {
        op[0] = mk<C>(e.x + o.x, 0);
        out_ny[line] = mk<C>(e.x - o.x, 0);
      } else {
        op[k] = cadd(e, cmul(twpl[k], o));
      }
    }
  }
  if constexpr (SAMPLE) block_sum<NTZ, ZPART>(acc, partial + (long long)wb * ZPART);
}

template <typename T, int H, int NTZ, bool FREE, int MODE, bool SAMPLE, bool PW>
__global__ __launch_bounds__(NTZ, (FREE ? 1 : (sizeof(T) == 4 ? 4 : 2)))
void k_zpass(const void* __restrict__ in, const cpx<T>* __restrict__ in_ny, void* __restrict__ out,
             cpx<T>* __restrict__ out_ny, long long nlines, ZLines zl, Batch bt,
             const cpx<T>* __restrict__ tw, const cpx<T>* __restrict__ twp,
             double* __restrict__ partial) {
  using C = cpx<T>;
  __shared__ C lds[z_lds_elems<T, H, NTZ>()];
  const int slot = bt.slots[blockIdx.y];
  if (bt.budget && bt.sib >= bt.budget[slot]) return;
  const long long so = (long long)slot * bt.stride;  // complex elements
  zpass_body<T, H, NTZ, MODE, SAMPLE, PW>(reinterpret_cast<const C*>(in) + so, in_ny ? in_ny + so : nullptr,
                                          reinterpret_cast<C*>(out) + so, out_ny ? out_ny + so : nullptr,
                                          nlines, zl, bt.ia[slot].za, tw, twp,
                                          partial + (long long)slot * bt.pzstride, lds, blockIdx.x);
}

}  // namespace xpfc
