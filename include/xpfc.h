/*
 * xpfc.h -- C ABI of the MI355X XPFC time-stepping library (libxpfc.so).
 *
 * The reference (LC2500/XPFC_CUDA) has no library: its operator API is the set
 * of free functions declared in globals.cuh:98-145 and driven by main.cu over
 * process globals (globals.cuh:36-96).  Each entry point below names the
 * reference function it replaces.  Differences that are deliberate:
 *   - state lives in an opaque context instead of globals;
 *   - errors are returned as status codes (never exit()), see xpfc_strerror;
 *   - results are returned to the caller instead of printed from the device
 *     (sum_reduction.cu:64,77);
 *   - fields are real arrays indexed [i][j] (2-D) or [i][j][l] (3-D), last
 *     index contiguous, exactly like psi_host[i*My + j] (initialize.cu:105);
 *   - spectra at the operator level use the real-transform half layout
 *     [n0][n1][n2/2+1] (the reference keeps full Z2Z arrays whose imaginary
 *     part of the inverse is discarded by copy_matrix.cu:44-45).
 * All pointers are plain host or device pointers; sizes are element counts.
 * A context is not re-entrant; distinct contexts may be used from distinct
 * host threads.
 */
#ifndef XPFC_H_
#define XPFC_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define XPFC_OK 0
#define XPFC_ERR_ARG 1         /* invalid argument or unsupported size      */
#define XPFC_ERR_IO 2          /* file could not be opened/read/written     */
#define XPFC_ERR_PARSE 3       /* malformed pfc.input or VTK                */
#define XPFC_ERR_DEVICE 4      /* HIP runtime error                         */
#define XPFC_ERR_NOMEM 5       /* device or host allocation failed          */
#define XPFC_ERR_STATE 6       /* call out of order (e.g. no field set)     */
#define XPFC_ERR_NONFINITE 7   /* a reduction produced NaN/Inf              */

#define XPFC_FP64 0            /* every field and transform in FP64          */
#define XPFC_FP32_MIXED 1      /* FP32 fields/FFTs; FP64 denominator, k=0, sums */

#define XPFC_DCF_SUM 0         /* initialize.cu:215-217 (reference active)   */
#define XPFC_DCF_MAX 1         /* initialize.cu:209-212 (envelope)           */

#define XPFC_MAX_PEAKS 4

/* Parameters.  Field names follow read_input.cu:19-115 and globals.cuh:40-52. */
typedef struct xpfc_params {
  int dim;                   /* 2 or 3 (the reference is 2-D, globals.cuh:26) */
  int n[3];                  /* Mx, My, Mz (Mz = 1 in 2-D): 2^k, 3*2^k or 5*2^k
                                points (k >= 2) in 8..2048; the reference takes any
                                Mx, My (read_input.cu:40-41) and plans cuFFT for them */
  double L[3];               /* Lx, Ly, Lz box lengths (ints in the reference) */
  double dt;                 /* time step                                       */
  double Mo;                 /* mobility                                        */
  double eps;                /* epsilon of the denominator (euler_step.cu:54)   */
  double a, b;               /* a = 1/a_in, b = 1/b_in (read_input.cu:28)       */
  double sigma;              /* s: DCF temperature parameter                    */
  double psi0;               /* mean density psi_0                              */
  int npeaks;                /* DCF peaks (initialize.cu:10-17 hard-codes 2)    */
  double peak_k[XPFC_MAX_PEAKS];
  double peak_alpha[XPFC_MAX_PEAKS];
  double peak_rho_beta[XPFC_MAX_PEAKS]; /* rho_i * B_i of exp(-(s k_i)^2/(2 rho_i B_i)) */
  int dcf_mode;              /* XPFC_DCF_SUM or XPFC_DCF_MAX                    */
  int precision;             /* XPFC_FP64 or XPFC_FP32_MIXED                    */
  int t_sample;              /* sampling period in steps (0 = never)           */
  int int_Mo_residual;       /* 1: residual uses (int)Mo (sse_calculation.cu:120) */
  /* pfc.input fields kept for the CLI and initial conditions */
  int no_val;
  int t_final;               /* t (int): the reference runs t/dt + 1 steps      */
  int iconfig;               /* 0 tri, 1 square, 2 liquid, 10 VTK restart       */
  double A_10, A_11, A_tri;
  int iseed, ishape;
  double h, k, roa;
  int device;                /* HIP device ordinal                              */
  int restart_order;         /* VTK restart mapping, see xpfc_inst_initialize_vtk */
  int powf_fp32;             /* 1: psi^2, psi^3 (and psi^4 of the free energy) formed
                                from float(psi) with FP32 products, as the reference's
                                powf (power_functions.cu:47); 0: exact products in the
                                working precision.  The reference's FP32 squaring of
                                the residual (sse_calculation.cu:38) is not reproduced:
                                the RMSE is a k-space FP64 sum. */
} xpfc_params;

/* One sampled diagnostic record (what sum_reduction.cu:64,77 prints). */
typedef struct xpfc_sample {
  long long step;            /* n of "iteration: n"                             */
  int has_fe;                /* FE of the field after step-1 (main.cu:412-426)  */
  int has_sse;               /* RMSE of the update n -> n+1 (main.cu:326-328)   */
  double F_ref;              /* ff_factor * sum f   (FE_calc.cu:64, initialize.cu:63) */
  double f_mean;             /* sum f / N                                        */
  double mean;               /* real-space sum psi / N (Avg_calc.cu:3-18)       */
  double rmse;               /* sqrt(sum R^2 / N)   (sum_reduction.cu:77)       */
} xpfc_sample;

typedef struct xpfc_ctx xpfc_ctx;

/* ---- parameters ---------------------------------------------------------- */
/* Defaults: the reference's hard-coded DCF (initialize.cu:10-17, sum mode),
 * a = 1, b = 1/3, Mo = 1, eps = 0, dt = 0.5, FP64. */
int xpfc_params_default(xpfc_params* out);
/* sizeof(xpfc_params) (no counterpart in the reference): a binding that
 * mirrors the struct checks its layout against it. */
size_t xpfc_params_size(void);
/* read_input (read_input.cu:8-121): parse a pfc.input file, same positional
 * format; a = 1/a_in, b = 1/b_in.  Optional "key value" lines after the
 * reference's lines (none in the reference's format): dim, Lz, Mz, precision
 * (fp64 | fp32), dcf (sum | max), peaks (n, then k alpha rho_beta per peak),
 * int_Mo_residual, powf_fp32, restart_order.  Returns XPFC_ERR_IO /
 * XPFC_ERR_PARSE. */
int xpfc_read_input(const char* path, xpfc_params* out);
/* N_t_steps = t/dt + 1 truncated as the loop bound uses it (initialize.cu:60). */
long long xpfc_reference_steps(const xpfc_params* p);

/* ---- context ------------------------------------------------------------- */
/* Allocates the device fields, twiddle tables and stream (main.cu:106-207). */
int xpfc_create(const xpfc_params* p, xpfc_ctx** out);
void xpfc_destroy(xpfc_ctx* ctx);
int xpfc_get_params(const xpfc_ctx* ctx, xpfc_params* out);

/* ---- initial conditions (initialize.cu:95-202, randn.cu:4-12) ------------ */
/* Reference initial condition from p->iconfig/iseed/ishape (2-D), with the
 * average-density fix of Avg_calc.cu:21-43; uploads the field. */
int xpfc_initialize(xpfc_ctx* ctx);
/* psi = psi0 + amplitude * randn(seed, cell), mean re-fixed to psi0. */
int xpfc_initialize_noise(xpfc_ctx* ctx, double amplitude, unsigned long long seed);
/* randn: count N(0,1) draws for cells first..first+count-1 (Philox4x32-10 +
 * Box-Muller, FP64).  Replaces the unseeded randn() of randn.cu:4-12. */
int xpfc_randn(unsigned long long seed, unsigned long long first, long long count, double* out);
/* Philox4x32-10 block function (exposed for known-answer tests). */
void xpfc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ---- state I/O (main.cu:226 H2D, copy_to_cpu.cu:3-18 D2H) ---------------- */
int xpfc_set_field(xpfc_ctx* ctx, const double* host_real);   /* N values */
int xpfc_get_field(xpfc_ctx* ctx, double* host_real);         /* N values */
long long xpfc_num_cells(const xpfc_ctx* ctx);

/* ---- the time-stepping hot path (main.cu:306-435) ------------------------ */
/* Advance nsteps steps.  Every t_sample steps the FE / RMSE records of the
 * reference schedule (SSE when n % t_sample == 0, FE when (n+1) % t_sample
 * == 0) are appended to the sample log.  The steps are queued on the context
 * stream and the call returns; the sampled sums are read back by xpfc_sync and
 * by the calls that read or change the sample log or an instance's state.  A
 * non-finite sum is kept in the record (NaN / Inf) and never stops the run of
 * any instance; xpfc_sync (and xpfc_time_steps) return XPFC_ERR_NONFINITE once
 * for each instance whose sampled sums became non-finite since the last such
 * report. */
int xpfc_step(xpfc_ctx* ctx, int nsteps);
long long xpfc_current_step(const xpfc_ctx* ctx);
/* Sample log: number of records, copy up to max of them (both return minus a
 * status code on failure), clear. */
int xpfc_num_samples(xpfc_ctx* ctx);
int xpfc_get_samples(xpfc_ctx* ctx, xpfc_sample* out, int max);
void xpfc_clear_samples(xpfc_ctx* ctx);
/* Time nsteps steps with hipEvents on the context stream (helpers_header.cuh:12-48). */
int xpfc_time_steps(xpfc_ctx* ctx, int nsteps, float* ms);
/* Per-pass timing: run nsteps unsampled steps with a hipEvent between every
 * launch; ms[0..3] = average ms of the axis-0 fused pass, the axis-1 inverse
 * pass, the real-space pass and the axis-1 forward pass (0 in 2-D), ms[4] =
 * average ms per whole step. */
int xpfc_time_passes(xpfc_ctx* ctx, int nsteps, double* ms);
/* The HBM rate of a plain streaming copy on `device`: a kernel that only
 * copies `bytes` (16 bytes per lane per access) between two fresh buffers,
 * timed with hipEvents over `reps` copies after one warm-up; *gbs = read +
 * written bytes per second / 1e9.  The practical ceiling the passes are
 * compared with (beside the 8 TB/s peak).  No counterpart in the reference. */
int xpfc_copy_bandwidth(int device, size_t bytes, int reps, double* gbs);
/* Pass variants: v[0] = tile width (columns per workgroup, 2/4/8/16) of the
 * axis-0 pass, v[1] = that of the axis-1 passes (plus 1: workgroups dealt to
 * the XCDs in runs of two tiles), v[2] = workgroup size (64..512) of the
 * real-space pass (plus 1: no occupancy hint).  Variants regroup lines into
 * workgroups, reorder workgroups or change register allocation, and never
 * change the computed field.  xpfc_autotune times every variant on scratch
 * buffers (two spectra per tuned instance when free device memory holds them,
 * else in place in the context's diagnostics buffer, so tuning allocates no
 * spectra the device cannot hold), keeps the fastest (cached per precision
 * and grid for the process)
 * and returns the choice in v (may be NULL); xpfc_step runs it on first use
 * unless XPFC_AUTOTUNE=0 or xpfc_set_variants was called (0 = keep). */
int xpfc_autotune(xpfc_ctx* ctx, int* v);
int xpfc_set_variants(xpfc_ctx* ctx, const int* v);
int xpfc_get_variants(const xpfc_ctx* ctx, int* v);
/* Step blocks: runs of unsampled steps are captured once as hipGraphs of up to
 * 50 steps and replayed (on by default; XPFC_GRAPHS=0 or on = 0 launches every
 * kernel eagerly).  Both give identical fields. */
int xpfc_set_graphs(xpfc_ctx* ctx, int on);
/* State form (chosen when a context is made; not an entry point): between
 * steps the state is held as psi_k, or, for 3-D grids whose spectrum is at
 * least 16 MiB, as psi_k transformed back along axis 0 (the axis-0 pass then
 * writes one spectrum less per step).  The two agree to the rounding of one
 * transform pair per step; XPFC_STATE_FORM=0 / 1 in the environment forces
 * psi_k / the axis-0 form.  Checkpoints record the form and load only into a
 * context of the same form. */
/* 1 when all work queued on the context has finished, 0 while it runs
 * (never blocks), minus a status code on failure. */
int xpfc_query(xpfc_ctx* ctx);
/* Divergence probe after the queued steps (blocks until they finish):
 * finite[s] = 1 while instance s's field is finite, 0 once it has diverged
 * (two spectral modes are tested; a non-finite value anywhere in the field
 * reaches every mode within one step).  finite[ninst]. */
int xpfc_probe_finite(xpfc_ctx* ctx, int* finite);
/* Block until all work of the context has finished; read back pending
 * samples.  Returns XPFC_ERR_NONFINITE (after completing the read-back) when
 * an instance's sampled sums became non-finite since the last report. */
int xpfc_sync(xpfc_ctx* ctx);

/* ---- diagnostics ----------------------------------------------------------- */
/* FE_calc_wrapper (FE_calc.cu:3-51) + convolution_wrapper (convolve_fields.cu:3-27)
 * for the current field: F_ref = ff_factor * sum f and f_mean = sum f / N. */
int xpfc_fe_calc(xpfc_ctx* ctx, double* F_ref, double* f_mean);
/* mean / Avg_calc (Avg_calc.cu:3-18): the FP64 mean density. */
int xpfc_avg_calc(xpfc_ctx* ctx, double* mean);
/* SSE_calc_wrapper (sse_calculation.cu:3-116): RMSE of the last sampled update. */
int xpfc_sse_calc(xpfc_ctx* ctx, double* rmse);

/* ---- operator level (device pointers; element type = context precision) ---- */
int xpfc_malloc(xpfc_ctx* ctx, void** dptr, size_t bytes);
int xpfc_free(xpfc_ctx* ctx, void* dptr);
int xpfc_memcpy_h2d(xpfc_ctx* ctx, void* dst, const void* src, size_t bytes);
int xpfc_memcpy_d2h(xpfc_ctx* ctx, void* dst, const void* src, size_t bytes);
size_t xpfc_real_bytes(const xpfc_ctx* ctx);      /* N reals                      */
size_t xpfc_spectrum_bytes(const xpfc_ctx* ctx);  /* n0*n1*(n2/2+1) complex       */
/* power_function_wrapper (power_functions.cu:4-53): dst = src^p on real fields
 * (exact products in the working precision, or FP32 products of float(src)
 * when the context's powf_fp32 is set). */
int xpfc_power_functions(xpfc_ctx* ctx, int p, const void* src, void* dst);
/* convolution_wrapper (convolve_fields.cu:3-51): out = c2(k) * psi_k. */
int xpfc_convolve_fields(xpfc_ctx* ctx, const void* psi_k, void* out);
/* euler_step_wrapper (euler_step.cu:3-63): out = semi-implicit update. */
int xpfc_euler_step(xpfc_ctx* ctx, const void* psi_k, const void* p2_k, const void* p3_k, void* out);
/* residual_calc (sse_calculation.cu:120-147): the k-space residual. */
int xpfc_residual_calc(xpfc_ctx* ctx, const void* psi_kp1, const void* psi_k, const void* p2_k,
                       const void* p3_k, void* out);
/* scaling_step_wrapper (scaling_step.cu:4-53): buf /= N (spectral=1: a
 * half spectrum; 0: a real field). */
int xpfc_scaling_step(xpfc_ctx* ctx, void* buf, int spectral);
/* cufftXtExecDescriptor FORWARD / INVERSE on real data (main.cu:309,343): the
 * unnormalised real-to-half-spectrum transform and its unnormalised inverse. */
int xpfc_fft_r2c(xpfc_ctx* ctx, const void* real_in, void* spec_out);
int xpfc_fft_c2r(xpfc_ctx* ctx, const void* spec_in, void* real_out);

/* ---- output (write_vtk.cu:4-35) ------------------------------------------- */
int xpfc_write_vtk(xpfc_ctx* ctx, const char* path);
/* Restart input (read_input.cu:124-241): read a legacy VTK structured-points
 * field (2-D or 3-D) into host memory, values in file order (x fastest);
 * dims[3] and spacing[3] returned; host_real may be NULL to query dims. */
int xpfc_read_vtk(const char* path, int* dims, double* spacing, double* host_real,
                  long long capacity);
/* Initial condition from a VTK file written by xpfc_write_vtk (or the
 * reference's writeVTKFile), with the average-density fix.  p->restart_order
 * selects how file values map to cells: 0 = x fastest (the file's own order;
 * write -> read is the identity), 1 = the reference's sequential fill
 * psi[i*My + j] (read_input.cu:208-218; transposes x and y, 2-D only).
 * xpfc_initialize with iconfig = 10 reads "initial_density_field.vtk" from the
 * working directory like the reference. */
int xpfc_inst_initialize_vtk(xpfc_ctx* ctx, int s, const char* path);
/* Binary checkpoint of one instance: parameters, step, FP64 mean and the
 * device buffers (the state in the context's form, psi_k or psi_k
 * transformed back along axis 0, and the nonlinear-term work field), so a run
 * resumed from it continues bit for bit.  Loading requires the same grid and
 * precision; a context of the other state form converts the state with one
 * axis-0 transform; the sample log is cleared. */
int xpfc_inst_save_checkpoint(xpfc_ctx* ctx, int s, const char* path);
int xpfc_inst_load_checkpoint(xpfc_ctx* ctx, int s, const char* path);
int xpfc_save_checkpoint(xpfc_ctx* ctx, const char* path);
int xpfc_load_checkpoint(xpfc_ctx* ctx, const char* path);

/* ---- several simulations of one grid in one context ------------------------ */
/* No counterpart in the reference, which runs one simulation per process
 * (main.cu:20-435).  Instance s of a context keeps its own physics (every
 * field of xpfc_params except the grid and precision may differ), field, step
 * counter and sample log; xpfc_step advances every ACTIVE instance with one
 * launch per pass.  An instance's field and samples are identical to those of
 * a one-instance context given the same parameters and field.  The functions
 * above act on instance 0. */
int xpfc_create_batch(const xpfc_params* p, int ninst, xpfc_ctx** out);  /* p[ninst] */
/* Device bytes xpfc_create_batch allocates per instance of this grid and
 * precision (0 for an unsupported grid): the unit of HBM admission. */
size_t xpfc_instance_bytes(const xpfc_params* p);
int xpfc_num_instances(const xpfc_ctx* ctx);
/* Active list (NULL = all instances, in order); stepping skips the others. */
int xpfc_set_active(xpfc_ctx* ctx, const int* inst, int n);
/* Replace instance s's physics (its field is kept; set one after this). */
int xpfc_inst_set_params(xpfc_ctx* ctx, int s, const xpfc_params* p);
/* Step counter and sample log back to zero (a slot handed to a new job). */
int xpfc_inst_reset(xpfc_ctx* ctx, int s);
int xpfc_inst_initialize(xpfc_ctx* ctx, int s);
int xpfc_inst_initialize_noise(xpfc_ctx* ctx, int s, double amplitude, unsigned long long seed);
/* xpfc_inst_initialize_noise without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (a serving loop that
 * admits jobs into resident slots without touching the device or host heap,
 * harnesses that forbid allocation).  The draws' partial sums go to a fixed
 * device array of 4096 doubles (32 KiB) that the library holds on each GPU
 * until the process ends; its device address is looked up at context
 * creation.  Calls on one GPU from several host threads share the array and
 * are serialised by the library.  The field is identical to
 * xpfc_inst_initialize_noise's.  Returns XPFC_ERR_STATE while sampled records
 * of instance s are pending (read them with xpfc_sync first), and
 * XPFC_ERR_ARG if the draws would need more partial sums than the array
 * holds.  On success it calls no hipMalloc / hipFree and no host allocator. */
int xpfc_inst_initialize_noise_noalloc(xpfc_ctx* ctx, int s, double amplitude, unsigned long long seed);
int xpfc_inst_set_field(xpfc_ctx* ctx, int s, const double* host_real);
int xpfc_inst_get_field(xpfc_ctx* ctx, int s, double* host_real);
long long xpfc_inst_current_step(const xpfc_ctx* ctx, int s);
int xpfc_inst_num_samples(xpfc_ctx* ctx, int s);
int xpfc_inst_get_samples(xpfc_ctx* ctx, int s, xpfc_sample* out, int max);
void xpfc_inst_clear_samples(xpfc_ctx* ctx, int s);
int xpfc_inst_fe_calc(xpfc_ctx* ctx, int s, double* F_ref, double* f_mean);
int xpfc_inst_avg_calc(xpfc_ctx* ctx, int s, double* mean);
int xpfc_inst_sse_calc(xpfc_ctx* ctx, int s, double* rmse);
int xpfc_inst_write_vtk(xpfc_ctx* ctx, int s, const char* path);

/* Parameters of instance s (any context). */
int xpfc_get_inst_params(const xpfc_ctx* ctx, int s, xpfc_params* out);

/* ---- simulations of several grids in one context --------------------------- */
/* No counterpart in the reference.  The instances may have different grids
 * (2-D and 3-D, extents of 2^k, 3*2^k or 5*2^k points up to 1024, any box
 * lengths; longer extents are refused with XPFC_ERR_ARG) but one precision
 * and device.  Every step advances all active instances of every grid with
 * one launch per pass and line-length family present (the axis-1 passes cover
 * the 3-D grids) in each of two stream groups (every grid's instances split
 * between them, a second stream forked from and joined back into the
 * context's; XPFC_MIXED_STREAMS=1 keeps one), and runs of unsampled steps are
 * replayed from captured hipGraphs as in one-grid contexts, so a block of
 * steps of the whole mix is one graph per group.  Each
 * instance's field is bitwise the one a one-instance context gives; sampled
 * FE / RMSE records sum their partials in another grouping (equal to
 * rounding), and xpfc_inst_fe_calc is bitwise the same.  The xpfc_inst_*
 * functions, xpfc_set_active (instance numbers of this context), xpfc_step,
 * xpfc_sync, xpfc_query, xpfc_probe_finite, xpfc_time_steps and the functions
 * on instance 0 accept it; the operator-level functions, xpfc_set_variants and
 * xpfc_time_passes refuse it (XPFC_ERR_ARG); xpfc_autotune has nothing to
 * choose (variants 0). */
int xpfc_create_mixed(const xpfc_params* p, int ninst, xpfc_ctx** out);  /* p[ninst] */
/* One block of len steps in which each active instance s takes steps[s]
 * (0..len) of them, all in one captured graph (any context: one grid, a
 * batch, or several grids): jobs of different remaining lengths advance
 * together without splitting the block.  No instance may have a sampled step
 * inside the block (XPFC_ERR_ARG; use xpfc_step for sampled runs).
 * steps[ninst].  No counterpart in the reference. */
int xpfc_step_budget(xpfc_ctx* ctx, int len, const int* steps);

/* ---- the job service ------------------------------------------------------- */
/* No counterpart in the reference.  Simulation jobs (a grid, physics, step
 * count and noise initial condition) are queued, admitted in arrival order
 * into free slots of their grid class while their device footprint
 * (xpfc_instance_bytes) fits the budget beside the resident jobs (a job that
 * does not fit blocks the later ones), and run in blocks of steps: the
 * classes of one lane advance together, every block of a lane advancing all
 * its running jobs with one launch per pass (a context of several grids), each
 * job by min(block, its remaining steps).  Lanes are independent (each its own
 * context and stream, with its own block cadence), so short jobs of a small
 * grid are not held to the blocks of a large one.  A job whose field diverges
 * stops at the end of that block.  Fields stay on the device
 * from admission to completion; a job's result (and kept field) is what the
 * same job gives alone in a one-instance context. */
#define XPFC_JOB_DONE 0
#define XPFC_JOB_NONFINITE 1      /* the field diverged; stopped early          */
#define XPFC_JOB_REJECTED 2       /* unknown grid class, steps < 1 or footprint > budget */

typedef struct xpfc_service xpfc_service;

typedef struct xpfc_service_class {
  int dim;                   /* 2 or 3                                          */
  int n[3];                  /* grid extents (n[2] ignored in 2-D)              */
  double L[3];               /* box lengths                                     */
  int slots;                 /* jobs of this grid that can run at once          */
  int lane;                  /* classes with equal lane numbers share blocks    */
} xpfc_service_class;

typedef struct xpfc_service_config {
  int precision;             /* XPFC_FP64 or XPFC_FP32_MIXED                    */
  int device;
  double budget_bytes;       /* device bytes the resident jobs may hold         */
  int block_steps;           /* steps per block while jobs wait (50)            */
  int lookahead;             /* blocks per launch while no job waits (1)        */
  int keep_fields;           /* 1: keep each finished job's field on the host   */
  int nclasses;
  const xpfc_service_class* classes;
} xpfc_service_config;

typedef struct xpfc_job {
  long long id;
  xpfc_params params;        /* grid (must match a class), physics; t_sample ignored */
  long long steps;
  double amplitude;          /* initial field psi0 + amplitude * randn(seed, cell) */
  unsigned long long seed;
  double arrival;            /* seconds on the service clock; < 0: now          */
} xpfc_job;

typedef struct xpfc_job_result {
  long long id;
  int status;                /* XPFC_JOB_*                                      */
  long long steps_done;
  double F_ref, f_mean, mean;  /* final free energy (xpfc_fe_calc) and mean     */
  double t_arrival, t_admitted, t_finished;  /* seconds on the service clock   */
} xpfc_job_result;

typedef struct xpfc_service_stats {
  long long submitted, rejected, done, nonfinite, queued, running, blocks, queue_max;
  double resident_bytes, peak_resident_bytes;
  double cell_updates;       /* cell updates run so far                         */
  double elapsed;            /* seconds since creation                          */
} xpfc_service_stats;

/* sizeof of the structs above (which = 0..4: class, config, job, result,
 * stats), for bindings that mirror them. */
size_t xpfc_service_struct_size(int which);
int xpfc_service_create(const xpfc_service_config* cfg, xpfc_service** out);
void xpfc_service_destroy(xpfc_service* svc);
/* Queue a job; XPFC_ERR_ARG (and a rejected result) if it cannot ever run. */
int xpfc_service_submit(xpfc_service* svc, const xpfc_job* job);
/* One non-blocking round (finish the lanes' blocks that are done, admit,
 * launch the next block of every lane between blocks): 1 if anything
 * happened, 0 if not, minus a status code. */
int xpfc_service_poll(xpfc_service* svc);
/* 1 when nothing is queued, running or in flight. */
int xpfc_service_idle(const xpfc_service* svc);
/* Serve `n` jobs: submit each at its arrival time (realtime = 1) or all at
 * once (0), poll until every job has finished.  Blocks. */
int xpfc_service_run(xpfc_service* svc, const xpfc_job* jobs, int n, int realtime);
/* Results of the jobs finished since the last call (returns the count copied,
 * at most max; the rest stay), minus a status code on failure. */
int xpfc_service_results(xpfc_service* svc, xpfc_job_result* out, int max);
/* The kept field of finished job `id` (keep_fields), capacity values;
 * returns the number of cells, minus a status code. */
long long xpfc_service_field(xpfc_service* svc, long long id, double* host, long long capacity);
int xpfc_service_get_stats(const xpfc_service* svc, xpfc_service_stats* out);

/* ---- errors ----------------------------------------------------------------- */
const char* xpfc_strerror(int status);
const char* xpfc_last_error(void);   /* detail of the last failure on this thread */

#ifdef __cplusplus
}
#endif
#endif /* XPFC_H_ */
