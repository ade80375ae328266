// xpfc_run: command-line driver with the reference's behaviour (main.cu:5-532)
// on one GPU: read pfc.input, build the initial condition, write
// density_data0.vtk, run N_t_steps = t/dt + 1 steps printing
// "iteration: n F = ... SSE_gpu = ..." on the reference's sampling schedule,
// print the TIMING line, write density_data<n>.vtk.
//
// usage: xpfc_run [pfc.input] [--steps N] [--noise AMP SEED] [--no-vtk] [--outdir DIR]
//                 [--resume CKPT] [--checkpoint CKPT]
// --resume continues a run saved with --checkpoint (binary device state, the
// run resumes bit for bit) instead of building the initial condition.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/xpfc.h"

static int die(int rc, const char* what) {
  std::fprintf(stderr, "xpfc_run: %s: %s (%s)\n", what, xpfc_strerror(rc), xpfc_last_error());
  return 1;
}

int main(int argc, char** argv) {
  const char* input = "pfc.input";
  long long steps = -1;
  bool noise = false, vtk = true;
  double amp = 0.0;
  unsigned long long seed = 0;
  std::string outdir = ".";
  const char* resume = nullptr;
  const char* ckpt = nullptr;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--steps") && i + 1 < argc) steps = std::atoll(argv[++i]);
    else if (!std::strcmp(argv[i], "--noise") && i + 2 < argc) {
      noise = true;
      amp = std::atof(argv[++i]);
      seed = std::strtoull(argv[++i], nullptr, 10);
    } else if (!std::strcmp(argv[i], "--no-vtk")) vtk = false;
    else if (!std::strcmp(argv[i], "--outdir") && i + 1 < argc) outdir = argv[++i];
    else if (!std::strcmp(argv[i], "--resume") && i + 1 < argc) resume = argv[++i];
    else if (!std::strcmp(argv[i], "--checkpoint") && i + 1 < argc) ckpt = argv[++i];
    else if (argv[i][0] != '-') input = argv[i];
    else {
      std::fprintf(stderr, "usage: xpfc_run [pfc.input] [--steps N] [--noise AMP SEED] [--no-vtk] "
                           "[--outdir DIR] [--resume CKPT] [--checkpoint CKPT]\n");
      return 2;
    }
  }
  xpfc_params p;
  int rc = xpfc_read_input(input, &p);
  if (rc) return die(rc, "read_input");
  if (steps < 0) steps = xpfc_reference_steps(&p);
  std::printf("Mx = %d My = %d Lx = %g Ly = %g dt = %g t_sample = %d steps = %lld\n", p.n[0], p.n[1],
              p.L[0], p.L[1], p.dt, p.t_sample, steps);
  xpfc_ctx* ctx = nullptr;
  if ((rc = xpfc_create(&p, &ctx))) return die(rc, "create");
  if (resume) rc = xpfc_load_checkpoint(ctx, resume);
  else rc = noise ? xpfc_initialize_noise(ctx, amp, seed) : xpfc_initialize(ctx);
  if (rc) return die(rc, resume ? "resume" : "initialize");
  if (vtk) {
    const std::string name = outdir + "/density_data" + std::to_string(xpfc_current_step(ctx)) + ".vtk";
    if ((rc = xpfc_write_vtk(ctx, name.c_str()))) return die(rc, "write_vtk");
  }
  float ms = 0.f;
  // steps in blocks so the sample log is printed as it grows
  long long done = 0;
  int printed = 0;
  std::vector<xpfc_sample> buf;
  const long long block = p.t_sample > 0 ? p.t_sample : 1000;
  while (done < steps) {
    const int nb = (int)((steps - done) < block ? (steps - done) : block);
    float t = 0.f;
    if ((rc = xpfc_time_steps(ctx, nb, &t))) return die(rc, "step");
    ms += t;
    done += nb;
    const int ns = xpfc_num_samples(ctx);
    if (ns < 0) return die(-ns, "samples");
    buf.resize(ns > 0 ? ns : 1);
    if ((rc = xpfc_get_samples(ctx, buf.data(), ns)) < 0) return die(-rc, "samples");
    // a record at step m is complete once step m itself has run (its RMSE
    // comes from the update m -> m+1); the last ones are printed after the loop
    const long long now = xpfc_current_step(ctx);
    for (; printed < ns && (buf[printed].step < now || done >= steps); ++printed) {
      const xpfc_sample& s = buf[printed];
      if (s.has_fe) std::printf("iteration: %lld F = %0.10f", s.step, s.F_ref);
      if (s.has_sse) std::printf(" SSE_gpu = %0.5e \n", s.rmse);
      else std::printf("\n");
    }
    std::fflush(stdout);
  }
  std::printf("TIMING: %g ms (Full iteration time.)\n", ms);
  double F = 0, fm = 0;
  if ((rc = xpfc_fe_calc(ctx, &F, &fm))) return die(rc, "fe_calc");
  std::printf("iteration: %lld F = %0.10f\n", xpfc_current_step(ctx), F);  // main.cu:450
  if (vtk) {
    const std::string name = outdir + "/density_data" + std::to_string(xpfc_current_step(ctx)) + ".vtk";
    if ((rc = xpfc_write_vtk(ctx, name.c_str()))) return die(rc, "write_vtk");
  }
  if (ckpt && (rc = xpfc_save_checkpoint(ctx, ckpt))) return die(rc, "checkpoint");
  xpfc_destroy(ctx);
  return 0;
}
