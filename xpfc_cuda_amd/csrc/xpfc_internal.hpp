// Internal structures shared by the host runtime and the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <mutex>

namespace xpfc {

// A family of equal-length lines visited by a strided FFT pass.  Line points
// are `rs` elements apart; neighbouring lines (columns) are `cs` apart; a
// batch of `ncols` columns repeats every `bs` elements.  `kind` tells the
// k-space update whether the lines belong to the main half spectrum (kind 0,
// column = kb*H + kz) or to the Nyquist plane kz = n2/2 (kind 1, column = kb).
struct Seg {
  long long off;
  long long bs;
  int rs, cs;
  int ncols, nbatch, ntc;
  int kind;
};
struct Segs {
  Seg s[2];
  int tiles0, tiles1;
  int lgc;  // workgroup order over the XCDs: runs of 2^lgc tiles, or < 0: one run per XCD
};

// Wavenumber geometry of the packed spectrum.  The main block is laid out
// [n0][n1][H] with planes (fixed axis-0 index) xs complex elements apart,
// xs >= n1*H (padding breaks the power-of-two plane stride); the Nyquist plane
// [n0][n1] starts at nyoff = n0*xs.
struct KGeom {
  int n0, n1, H;
  long long xs, nyoff;
  double dk0, dk1, dk2;  // 2*pi/L per axis
  // Equal spacing on every axis (square / cubic boxes): k^2 = dksq * m with the
  // integer m = i^2 + j^2 + l^2, and C2 comes from a per-instance table over m.
  int cubic;
  double dksq;
  // State form: 0 = psi_k; 1 = psi_k transformed back along axis 0 (then
  // psi_k = FFT_0(state) * inv_n0, and the axis-0 pass does not write psi_k).
  int xreal;
  double inv_n0;  // 1 / n0
};

// Physics constants of the semi-implicit update (all FP64).
// One row of the C2 table of a 2-D box that is not square: the modes
// |i| = 0 .. len-1 of column kz start at entry off.
struct C2Row {
  int off, len;
};

struct Phys {
  double dtM;        // dt * Mo
  double one_m_eps;  // 1 - eps
  double inv_dt;     // 1 / dt          (residual)
  double Mres;       // residual mobility (Mo, or (int)Mo like the reference)
  double ecut;       // peak terms exp(-e) with e >= ecut are dropped (below the
                     // working precision of the denominator: 1e-10 FP32, 4e-18 FP64)
  double kcut2;      // |k|^2 beyond which every peak term has e >= ecut
  int npk;
  int dcf_max;       // 0 = sum of peaks, 1 = max over peaks
  double pk[4], inv2a2[4], amp[4];
  // cubic geometry: c2tab[m] = C2(sqrt(dksq * m)) for m < mcut, 0 beyond
  int mcut;
  const double* c2tab;
  // 2-D box that is not square: c2tab[c2rows[kz].off + |i|] for kz < c2nl and
  // |i| < c2rows[kz].len, 0 beyond (null: evaluated per mode)
  const C2Row* c2rows;
  int c2nl;
};

// Real-space pass constants.
struct ZArgs {
  double mean;   // FP64 mean density (the k = 0 mode)
  double a, b;   // N(psi) = -a psi^2 + b psi^3
  double scale;  // 1/N for the inverse transform
  int powf32;    // powers from float(psi) with FP32 products (power_functions.cu:47)
};

enum ZMode { Z_STEP = 0, Z_R2C_FLUCT = 1, Z_R2C_NL = 2, Z_C2R = 3 };

// Line placement of the real-space pass: line l = a*n1 + b starts at
// a*xs + b*H complex elements (input and output may use different plane
// strides, e.g. a caller's contiguous real field and a padded spectrum).
struct ZLines {
  int n1;
  long long xs_in, xs_out;
};

// Per-instance constants of a context holding several simulations of one grid
// (device array indexed by instance slot).
struct InstArgs {
  Phys ph;
  ZArgs za;
};

// Which instances a launch covers: blockIdx.y (blockIdx.x for the finalize)
// indexes `slots`; instance `s` has its spectra `s * stride` complex elements
// from the base pointers, its X partials `s * pstride` and its real-space
// partials `s * pzstride` doubles from the partial base pointers.
// In a budgeted block (xpfc_step_budget) the launches of step `sib` of the
// block carry `budget`: the workgroups of an instance whose budget[slot] <= sib
// return at once, so each instance takes its own number of the block's steps.
struct Batch {
  const int* slots;
  const InstArgs* ia;
  long long stride;
  long long pstride;
  long long pzstride;
  const int* budget;  // nullptr: every listed instance steps
  int sib;
};
// Register-allocation variant flag (the arithmetic is identical), or-ed into
// the real-space workgroup size: no occupancy hint.
constexpr int ZFREE_REGS = 1;
// Bit 0 of an axis-1 pass variant (tile width 2, 4, 8 or 16): workgroups dealt to
// the XCDs in runs of two tiles instead of one contiguous run per XCD.
constexpr int YRUNS_PAIRED = 1;

// Per-workgroup FP64 partial sums.  X pass: {sum w c2 |psi_k|^2, sum w |R_k|^2}.
// Z pass: {sum psi, sum psi^2, sum psi^3, sum psi^4}.
constexpr int XPART = 2;
constexpr int ZPART = 4;
constexpr int NSUMS = 6;

// ---------------------------------------------------------------------------
// Ragged launches: one launch per pass over the instances of several grids
// (size classes).  Workgroups [wg0, wg0 + nact * wgs) of the launch belong to
// class k, instance-major; a workgroup finds its class by scanning wg0.  The
// kernels are compiled per line-length family (2^k, 3*2^k, 5*2^k points), so
// a pass over classes of several families is one launch per family.  The
// workgroup size is fixed per family and pass (rag_nt for the strided passes,
// RAG_NTZ for the real-space pass), so a class with lines of L points takes
// tiles of rag_nt * E / L columns (E points per thread) and RAG_NTZ /
// (threads per line) lines; the arithmetic of every line is that of the
// one-grid kernels.
// ---------------------------------------------------------------------------
constexpr int RAG_MAX = 16;    // classes per ragged launch
constexpr int RAG_NTZ = 256;   // threads per workgroup, real-space pass
struct RClass {
  int L;            // line length of the pass (n0, n1, or H = n2 / 2)
  int wg0;          // first workgroup of the class
  int wgs;          // workgroups per instance
  int nact;         // instances of the class in this launch
  int cubic;        // axis-0 pass: C2 from the table over m
  int pw;           // real-space pass: some instance forms the FP32 powers
  int lgc;          // strided passes: workgroup order over the XCDs within the class
                    // (runs of 2^lgc tiles, or < 0: one run per XCD); wg0 is a
                    // multiple of 8, so a workgroup's XCD is its index in the class mod 8
  Segs segs;        // strided passes: the class's segments, tiles filled
  KGeom g;          // axis-0 pass
  Batch bt;         // slots (the class's active list), instance table, strides
  void* a;          // axis-0: work; axis-1 and real space: the output (work)
  void* b;          // axis-0: psi (the state); axis-1: the inverse pass's input;
                    // real space: the input (psi only for 2-D grids in the axis-0 state form)
  const void* tw;   // line twiddles
  const void* twp;  // real space: half-spectrum twiddles
  double* partial;  // sampled partial sums (X or Z) of the class
  long long nlines; // real space: lines per instance
  ZLines zl;        // real space
  long long nyoff;  // real space: offset of the Nyquist plane (complex elements)
  const int* budget;  // steps of the current block per slot (nullptr: every step)
};
// Family of a line length: 1 (2^k), 3 (3*2^k) or 5 (5*2^k points).
constexpr int line_family(int L) { return L % 5 == 0 ? 5 : L % 3 == 0 ? 3 : 1; }
// Points per thread of a strided line (the FFT's per-thread radix budget).
constexpr int line_elems(int L) {
  return L % 5 == 0 ? (L < 20 ? L : 20) : L % 3 == 0 ? (L < 12 ? L : 12) : (L < 8 ? L : 8);
}
// Threads per workgroup of a family's ragged strided passes.  12- and
// 20-point lines take 256 (one wave per SIMD, so the axis-0 pass may hold up
// to 512 registers per lane: at 512 threads its FP64 form spills; 20-point FP64
// tiles of 512 threads would also not fit the LDS).
constexpr int rag_nt(int fam) { return fam == 1 ? 512 : 256; }
// Column tile width of a ragged strided pass over lines of L points.
constexpr int rag_tile(int L) {
  return L >= 8 ? rag_nt(line_family(L)) * line_elems(L) / L : 0;
}
// The ragged launchers of one precision and family (xpfc_ragged.hpp,
// instantiated by xpfc_rag_*.hip).
template <typename T, int FAM>
struct RagOps {
  static int prepare();
  static int xstep(const RClass* cls, int ncls, int nwg, size_t lds, bool lut, bool sample, int sib,
                   hipStream_t s);
  static int strided(const RClass* cls, int ncls, int nwg, size_t lds, int dir, int sib, hipStream_t s);
  static int zpass(const RClass* cls, int ncls, int nwg, size_t lds, bool sample, bool pw, int sib,
                   hipStream_t s);
  static int strided_wgs(int L, const Segs& segs);
  static long long zpass_wgs(int H, long long nlines, size_t* lds);
};

// Launch entry points of the step passes (xpfc_passes.hip).  `tc` is the tile
// width (2, 4, 8 or 16 columns per workgroup) of a strided pass and `ntz` the
// workgroup size (64..512) of the real-space pass; a combination a size does
// not support returns -3.  launch_xstep returns its workgroup count (= number
// of X partials) on success.
// Launchers return launch_status(v): v when the launches just made succeeded,
// else -2, with the HIP error kept for the caller's message (launch_error(),
// per host thread).
int launch_status(int ok_value = 0);
hipError_t launch_error();

// `ny` is the number of instances covered (entries of bt.slots used).  `nt`
// selects the streaming-hint loads and stores of the strided pass, which
// reads `in` and writes `buf` (the same buffer, or two of one layout).
int launch_strided(int prec, const void* in, void* buf, const Segs& segs, int L, int tc, int dir, const void* tw,
                   const Batch& bt, int ny, bool nt, hipStream_t s);
inline int launch_strided(int prec, void* buf, const Segs& segs, int L, int tc, int dir, const void* tw,
                          const Batch& bt, int ny, bool nt, hipStream_t s) {
  return launch_strided(prec, buf, buf, segs, L, tc, dir, tw, bt, ny, nt, s);
}
int launch_xstep(int prec, void* work, void* psi, const Segs& segs, int L, int tc, const void* tw,
                 const KGeom& g, const Batch& bt, int ny, double* partial, bool sample,
                 hipStream_t s);
int launch_zpass(int prec, int mode, const void* in, const void* in_ny, void* out, void* out_ny,
                 long long nlines, int H, const ZLines& zl, int ntz, const Batch& bt, int ny,
                 const void* tw, const void* twp, double* partial, bool sample, bool pw,
                 hipStream_t s);
int zpass_blocks(int prec, int H, int ntz, long long nlines);
// Ragged launches over the device table `cls` (ncls classes of line family
// `fam`, `nwg` workgroups in all, `lds` bytes of dynamic LDS).  `sib`: index
// of the step within its block, compared with the per-slot budgets.  Return
// 0, or -2 / -3 as above.  (`lut`: every class of this table has a square /
// cubic box)
int launch_xstep_ragged(int prec, int fam, const RClass* cls, int ncls, int nwg, size_t lds, bool lut,
                        bool sample, int sib, hipStream_t s);
int launch_strided_ragged(int prec, int fam, const RClass* cls, int ncls, int nwg, size_t lds, int dir,
                          int sib, hipStream_t s);
int launch_zpass_ragged(int prec, int fam, const RClass* cls, int ncls, int nwg, size_t lds, bool sample,
                        bool pw, int sib, hipStream_t s);
// Lets the ragged kernels of precision prec and family fam use the dynamic
// LDS of the family's longest tiles (call before launching or capturing them).
int rag_prepare(int prec, int fam);
// Workgroups per instance and dynamic LDS bytes of a class in a ragged launch
// (0 / negative: lines of that length have no ragged form).
int rag_strided_wgs(int prec, int L, const Segs& segs);
size_t rag_strided_lds(int prec, int L);
int rag_zpass_wgs(int prec, int H, long long nlines);
size_t rag_zpass_lds(int prec, int H);
int strided_blocks(int prec, int L, int tc, const Segs& segs);
// out[slot * NSUMS + k]: the six sums of every listed instance
int launch_finalize(const double* px, int nx, const double* pz, int nz, const Batch& bt, int ny,
                    double* out, hipStream_t s);
int launch_kspace_sums(int prec, const void* psi, long long nmain, long long nny, const KGeom& g,
                       const Batch& bt, double* partial, int* nblocks, hipStream_t s);
// tab[m] = C2(sqrt(dksq * m)), m < mcut, for the physics `ph`
int launch_c2table(double* tab, int mcut, double dksq, const Phys& ph, hipStream_t s);
int launch_c2rows(double* tab, const C2Row* rows, int nl, double dk0, double dk2, const Phys& ph, hipStream_t s);

// operator-level kernels on natural layouts (real [N], spectrum [n0][n1][n2/2+1])
int launch_power(int prec, const void* src, void* dst, long long n, int p, bool f32, hipStream_t s);
int launch_scale(int prec, void* buf, long long ncomplex_or_real, double f, hipStream_t s);
int launch_convolve(int prec, const void* psi_k, void* out, const KGeom& g, const Phys& ph,
                    int n2, hipStream_t s);
int launch_euler(int prec, const void* psi_k, const void* p2_k, const void* p3_k, void* out,
                 const KGeom& g, const Phys& ph, double a, double b, int n2, hipStream_t s);
int launch_residual(int prec, const void* psi_kp1, const void* psi_k, const void* p2_k,
                    const void* p3_k, void* out, const KGeom& g, const Phys& ph, double a,
                    double b, int n2, hipStream_t s);
// natural [n0][n1][H+1] <-> packed main/Nyquist layout (plane stride xs)
int launch_pack(int prec, const void* natural, void* main_ny, int n0n1, int H, int n1, long long xs,
                hipStream_t s);
int launch_unpack(int prec, const void* main_ny, void* natural, int n0n1, int H, int n1, long long xs,
                  hipStream_t s);
int launch_to_double(int prec, const void* src, double* dst, long long n, hipStream_t s);
// noise initial condition: per-workgroup sums of the draws of cells 0..n-1
// (nblocks returned, <= 4096), then out[c] = amp * (z_c - mean_z) as reals
constexpr int NOISE_PARTIALS = 4096;  // most per-workgroup sums of launch_noise_sum
int noise_partials(long long n);     // the number launch_noise_sum produces for n cells
int launch_noise_sum(long long n, unsigned long long seed, double* partial, int* nblocks, hipStream_t s);
// (the reals of cell c land at (c / plane) * pitch + c % plane)
int launch_noise_fill(int prec, void* out, long long n, long long plane, long long pitch,
                      unsigned long long seed, double amp, double mean_z, hipStream_t s);
// flag[s] = 1 if instance s's spectrum (at s * stride) is finite at two probe modes
// plain streaming copy of `bytes` (multiple of 16) from src to dst
int launch_stream_copy(const void* src, void* dst, size_t bytes, hipStream_t s);
int launch_probe(int prec, const void* psi, long long stride, long long iny, int K, int* flag,
                 hipStream_t s);
int launch_from_double(int prec, const double* src, void* dst, long long n, hipStream_t s);
// Fixed device array (NOISE_PARTIALS doubles per GPU, held until the process
// ends) for the allocation-free noise initialisation.  setup: looks up the
// array's device address with `device` current (once per device; 0, or < 0
// on failure); noalloc_scratch: that address (nullptr before setup); the lock
// serialises the host threads that use one device's array.
int noalloc_scratch_setup(int device);
double* noalloc_scratch(int device);
std::mutex* noalloc_scratch_lock(int device);

}  // namespace xpfc
