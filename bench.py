"""XPFC time-stepping throughput on one MI355X.

Workload: 3-D fcc two-peak XPFC, 512^3 periodic grid, dx = 0.125, dt = 0.5,
peaks k1 = 2π, k2 = 4π/√3 (α = 1, σ = 0.1, max envelope), n0 = 0.05 + 0.01 N(0,1)
(seed 2), FP32 fields and FFTs with FP64 denominator / k = 0 mode / reductions.
A step is one full semi-implicit update of the field resident in HBM (the
reference's loop body, main.cu:306-435); FE/RMSE are sampled every 100 steps.

Prints one JSON line.  `value` = steps/s of the FP32-mixed path; the FP64
plain path at the same size is timed beside it (`fp64_steps_per_s`).  Two more
workloads ride along in their own keys:
  batched -- 64 independent 2-D square-lattice instances, sides drawn from
             {128, 256, 512, 1024}, sigma ~ U[0.05, 0.2], n0 ~ U[0.02, 0.08],
             noise amplitude 0.01 seeded by the instance id, FP32-mixed, 2000
             steps each (FE/RMSE every 100) in 50-step graph blocks;
  serving -- Poisson arrivals of 2-D (128^2-1024^2) and 3-D (64^3-256^3) jobs,
             sides log-uniform, 500-5000 steps, admitted under a 200 GB budget;
             p50/p99 time-to-completion and aggregate cell-updates/s.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# The serving leg runs one stream per size class (17 classes with 2^k, 3*2^k
# and 5*2^k sides).  With HIP's default of 4 hardware queues per process those
# streams share queues and a small grid's block waits behind a large grid's
# (measured, 30 s stream: p50 1.52 s with 4 queues, 0.64 s with 32).  The HIP
# runtime reads the count when it starts, so it is raised (never lowered)
# before anything touches the GPU.
HW_QUEUES = 32
try:
    if int(os.environ.get("GPU_MAX_HW_QUEUES", "0")) < HW_QUEUES:
        os.environ["GPU_MAX_HW_QUEUES"] = str(HW_QUEUES)
except ValueError:
    os.environ["GPU_MAX_HW_QUEUES"] = str(HW_QUEUES)

HBM_PEAK_GBS = 8000.0
# Bytes in spectra S (one packed half spectrum).  The compulsory-traffic model
# of the step (one read and one write per axis pass, the state kept as psi_k)
# gives the axis-0 pass [FFT + update + inverse FFT] 4 S and a 3-D step 10 S:
# the algorithmic bytes of the roofline.  The 512^3 workload keeps the state
# along axis 0 in real space, which drops one write: its axis-0 pass moves
# 3 S (the measured traffic) and its step 9 S.
XSTEP_MODEL_SPECTRA = 4
STEP_MODEL_SPECTRA_3D = 10
XSTEP_SPECTRA = 3
STEP_SPECTRA_3D = 9


def fcc_params(n, prec, t_sample=100):
    import xpfc_cuda_amd as X
    k1 = 2.0 * math.pi
    k2 = k1 * 2.0 / math.sqrt(3.0)
    L = n * 0.125
    return X.Params(n=(n, n, n), L=(L, L, L), dt=0.5, Mo=1.0, eps=0.0, a=1.0, b=1.0 / 3.0,
                    sigma=0.1, psi0=0.05, peaks=((k1, 1.0, 1.0), (k2, 1.0, 1.0)),
                    dcf_mode=X.XPFC_DCF_MAX, precision=prec, t_sample=t_sample)


def spectrum_bytes(n, prec):
    # packed half spectrum: n*n*(n/2) main + n*n Nyquist plane, complex
    csize = 8 if prec == 1 else 16
    return n * n * (n // 2 + 1) * csize


def copy_ceiling(nbytes):
    import xpfc_cuda_amd as X
    return max(X.copy_bandwidth(nbytes, reps=20) for _ in range(3))


def gpu_run(n, prec, steps, warmup, seed=2, pin=None):
    import xpfc_cuda_amd as X
    p = fcc_params(n, prec)
    with X.Context(p) as ctx:
        if pin:
            ctx.set_variants(*pin)   # no autotuning
        ctx.initialize_noise(0.01, seed)
        if warmup:
            ctx.step(warmup)
        ctx.sync()
        t0 = time.perf_counter()
        ms = ctx.time_steps(steps)
        ctx.sync()
        wall = time.perf_counter() - t0
        passes = ctx.time_passes(max(5, min(steps, 50)))
        samples = ctx.samples()
        variants = ctx.variants()
        # host-buffer boundary (outside the timed region): field in and out
        io = None
        if prec == 1:
            t0 = time.perf_counter()
            field = ctx.get_field()
            t1 = time.perf_counter()
            ctx.set_field(field)
            t2 = time.perf_counter()
            io = {"get_field_ms": round((t1 - t0) * 1e3, 2), "set_field_ms": round((t2 - t1) * 1e3, 2)}
    return ms, wall, passes, samples, variants, io


def accuracy_run(n, steps=1000, seed=2):
    """FP32-mixed against the FP64 path on the same workload after `steps`
    steps: max |dn| and the relative free-energy difference (the FP64 path is
    parity-tested against the FP64 oracle; the oracle itself is too slow at
    this size)."""
    import numpy as np
    import xpfc_cuda_amd as X
    fields, fe = [], []
    for prec in (1, 0):
        with X.Context(fcc_params(n, prec, t_sample=0)) as ctx:
            ctx.initialize_noise(0.01, seed)
            ctx.step(steps)
            fields.append(ctx.get_field())
            fe.append(ctx.fe_calc()[0])
    dn = float(np.max(np.abs(fields[0] - fields[1])))
    return {"steps": steps, "max_abs_dn": dn, "fe_rel": abs(fe[0] - fe[1]) / abs(fe[1]),
            "F_fp32": fe[0], "F_fp64": fe[1],
            "reference": "FP64 GPU path on the same initial field"}


def _oracle_fcc(O, n, sigma, t_sample=100):
    k1 = 2.0 * math.pi
    k2 = k1 * 2.0 / math.sqrt(3.0)
    L = n * 0.125
    return O.Params(n=(n, n, n), L=(L, L, L), sigma=sigma, psi0=0.05,
                    peaks=[O.Peak(k1, 1.0, 1.0), O.Peak(k2, 1.0, 1.0)], dcf_mode="max",
                    t_sample=t_sample)


def cpu_baseline(n, nsteps_cap=6, budget_s=20.0):
    """The CPU leg: the FP64 oracle (numpy/scipy pocketfft, the host share of
    threads) timed on a bounded number of steps of the same 512^3 workload,
    started from the GPU's own initial field; returns the baseline record and
    the oracle's field after `steps` steps, the checker of the GPU run."""
    import numpy as np
    import xpfc_cuda_amd as X
    from oracle import xpfc_oracle as O
    # the host share of one GPU is 16 threads; never more than the machine has
    workers = max(1, min(int(os.environ.get("OMP_NUM_THREADS", "16")), os.cpu_count() or 1))
    op = _oracle_fcc(O, n, 0.1)
    with X.Context(fcc_params(n, 1)) as ctx:
        ctx.initialize_noise(0.01, 2)
        psi = ctx.get_field()
    st = O.Stepper(psi, op, workers)   # tables built outside the timed window
    st.step()                          # warm plans and pages
    done = 0
    t0 = time.perf_counter()
    while done < nsteps_cap and time.perf_counter() - t0 < budget_s:
        st.step()
        done += 1
    dt = time.perf_counter() - t0
    rec = {"value": done / dt, "unit": "steps/s", "cores": workers, "kind": "port",
           "sample": "%d step(s) of the %d^3 FP64 oracle (scipy.fft, %d workers) on %s"
                     % (done, n, workers, _cpu_model()),
           "host_cpus": os.cpu_count()}
    return rec, psi, st.psi, done + 1, workers


def accuracy_vs_oracle(n, psi0, ref, nsteps, workers):
    """The GPU against the FP64 oracle (the checker; bench.py's CPU leg only):
    (a) the benchmark workload itself after the oracle's `nsteps` steps, from
    the oracle's initial field psi0, FP32-mixed and FP64; (b) the same fcc physics made
    crystallising (sigma 0.07) at 128^3, 116 steps, where the amplitude passes
    0.1 (the nonlinear term matters): max |dn|, relative to the amplitude, and
    the relative free-energy difference at the last sample."""
    import numpy as np
    import xpfc_cuda_amd as X
    from oracle import xpfc_oracle as O
    out = {}
    at = {"steps": nsteps}
    for prec, key in ((1, "fp32_mixed"), (0, "fp64")):
        with X.Context(fcc_params(n, prec, t_sample=0)) as ctx:
            ctx.set_field(psi0)
            ctx.step(nsteps)
            got = ctx.get_field()
        at[key] = {"max_abs_dn": float(np.max(np.abs(got - ref)))}
        del got
    out["workload_%dcube" % n] = at
    m, steps = 128, 116
    op = _oracle_fcc(O, m, 0.07, t_sample=20)
    p32 = fcc_params(m, 1, t_sample=20)
    p32.sigma = 0.07
    with X.Context(p32) as ctx:
        ctx.initialize_noise(0.01, 2)
        psi0 = ctx.get_field()
    oref, osm = O.run(psi0, op, steps, sample=True, workers=workers)
    amp = float(np.max(np.abs(oref - 0.05)))
    f_or = [s.value for s in osm if s.kind == "FE" and s.step == steps][-1]
    cr = {"steps": steps, "amplitude": amp}
    for prec, key in ((1, "fp32_mixed"), (0, "fp64")):
        p = fcc_params(m, prec, t_sample=20)
        p.sigma = 0.07
        with X.Context(p) as ctx:
            ctx.set_field(psi0)
            ctx.step(steps)
            got = ctx.get_field()
            F, _ = ctx.fe_calc()
        dn = float(np.max(np.abs(got - oref)))
        cr[key] = {"max_abs_dn": dn, "rel_to_amplitude": dn / amp, "fe_rel": abs(F - f_or) / abs(f_or)}
    out["crystallising_128cube_sigma0.07"] = cr
    out["reference"] = "FP64 oracle (numpy/scipy) on the same initial field"
    return out


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def config_runs():
    """The other single-simulation configurations: C1 (2-D 256^2 square lattice,
    FP64, 1000 steps, FE/RMSE every 100) and C2 (3-D 256^3 bcc, FP64, 1000
    steps), timed with hipEvents after 20 warm-up steps."""
    import xpfc_cuda_amd as X
    k1 = 2.0 * math.pi
    out = {}
    cases = {
        "c1_2d_256_fp64_square": dict(n=(256, 256), L=(32.0, 32.0), sigma=0.1, psi0=0.05, seed=0,
                                      peaks=((k1, 1.0, 1.0), (k1 * math.sqrt(2.0), 1.0, 1.0))),
        "c2_3d_256_fp64_bcc": dict(n=(256, 256, 256), L=(32.0, 32.0, 32.0), sigma=0.12, psi0=0.04,
                                   seed=1, peaks=((k1, 1.0, 1.0),)),
    }
    for name, cfg in cases.items():
        p = X.Params(n=cfg["n"], L=cfg["L"], dt=0.5, Mo=1.0, eps=0.0, a=1.0, b=1.0 / 3.0,
                     sigma=cfg["sigma"], psi0=cfg["psi0"], peaks=cfg["peaks"],
                     dcf_mode=X.XPFC_DCF_MAX, precision=X.XPFC_FP64, t_sample=100)
        with X.Context(p) as ctx:
            ctx.initialize_noise(0.01, cfg["seed"])
            ctx.step(20)
            ctx.sync()
            ms = ctx.time_steps(1000)
            cells = 1
            for m in cfg["n"]:
                cells *= m
            out[name] = {"steps_per_s": round(1000.0 / (ms * 1e-3), 1),
                         "ms_per_step": round(ms / 1000.0, 5),
                         "cell_updates_per_s": cells * 1000.0 / (ms * 1e-3),
                         "last_sample": ctx.samples()[-1].line().strip()}
    return out


def batched_run(steps=2000, ninst=64, seed=0):
    """The 64-instance batch of 2-D grids of four sizes in one mixed-grid
    context: every step is one launch per pass over the instances of each of
    two stream groups (every grid's instances split between them, so their
    launches overlap), and unsampled runs replay 50-step graphs of those
    launches (FE / RMSE every 100 steps).  Wall time of the whole job, and beside it the same job as one
    context per grid size on its own stream (the classes overlap on the GPU)."""
    import xpfc_cuda_amd as X
    from xpfc_cuda_amd import serving as S
    draws = S.batch_jobs(ninst, steps, seed)
    ps = []
    for j in draws:
        p = S.job_params(j)
        p.t_sample = 100
        ps.append(p)
    cells = sum(j.cells for j in draws)

    def finish(sync):
        try:
            sync()
        except X.XPFCError as e:
            # strongly crystallising draws (small sigma, large n0) diverge under
            # the explicit nonlinear term, as in the FP64 oracle; counted below
            if e.status != X.api.XPFC_ERR_NONFINITE:
                raise

    mc = X.MixedContext(ps)
    for s, j in enumerate(draws):
        mc.initialize_noise(s, j.amplitude, j.seed)
    mc.sync()
    t0 = time.perf_counter()
    mc.step(steps)
    finish(mc.sync)
    wall = time.perf_counter() - t0
    diverged = 0
    fe = {}
    for s, j in enumerate(draws):
        smp = mc.samples(s)
        diverged += not all(math.isfinite(r.F_ref) for r in smp if r.has_fe)
        if str(j.side) not in fe:
            fe[str(j.side)] = [smp[-1].step, smp[-1].F_ref]
    mc.close()
    res = {
        "instances": ninst, "steps": steps, "wall_s": round(wall, 4),
        "instance_steps_per_s": round(ninst * steps / wall, 1),
        "cell_updates_per_s": cells * steps / wall,
        "launches": "one per pass over all grids in each of two stream groups (mixed-grid context; every "
                    "grid's instances split between the groups), 50-step graphs per group",
        "sides": {str(side): sum(j.side == side for j in draws) for side in sorted({j.side for j in draws})},
        "last_fe_first_instance": fe, "nonfinite_instances": diverged,
    }
    # the same job as one context per grid size, each on its own stream
    groups = {}
    for j in draws:
        groups.setdefault(j.side, []).append(j.jid)
    ctxs = []
    for side, ids in sorted(groups.items()):
        bc = X.BatchContext([ps[i] for i in ids])
        bc.autotune()
        for s, i in enumerate(ids):
            bc.initialize_noise(s, draws[i].amplitude, draws[i].seed)
        ctxs.append(bc)
    for bc in ctxs:
        bc.sync()
    t0 = time.perf_counter()
    for bc in ctxs:
        bc.step(steps)          # queued: the classes overlap on their streams
    for bc in ctxs:
        finish(bc.sync)
    wall2 = time.perf_counter() - t0
    for bc in ctxs:
        bc.close()
    res["per_size_contexts"] = {"wall_s": round(wall2, 4), "cell_updates_per_s": cells * steps / wall2}
    return res


def serving_run(seconds, rate, frac_3d=0.5, seed=0, budget=200e9, class_cells=1 << 25,
                lookahead=1, lanes="class", rounding="smooth"):
    """Poisson job stream through the continuous-batching service."""
    from xpfc_cuda_amd import serving as S
    jobs = S.poisson_jobs(rate, seconds, seed=seed, frac_3d=frac_3d, rounding=rounding)
    svc = S.Service(budget_bytes=budget, slots=S.class_slots(class_cells), lookahead=lookahead,
                    lanes=S.lane_groups(lanes))
    try:
        svc.warm([(2, v) for v in S.grid_sides(128, 1024, rounding)] +
                 [(3, v) for v in S.grid_sides(64, 256, rounding)])
        rep = svc.run(jobs, realtime=True,
                      progress=lambda m: print("serving " + m, file=sys.stderr, flush=True))
        # offered load: cell-updates per second the arrivals ask for
        offered = sum(j.cells * j.steps for j in jobs) / seconds
        rep.update({"arrival_rate": rate, "window_s": seconds, "frac_3d": frac_3d,
                    "offered_cell_updates_per_s": offered,
                    "budget_gb": budget / 1e9, "lookahead_blocks": lookahead, "lanes": lanes,
                    "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "0")),
                    "rounding": ("side log-uniform, rounded to the nearest 2^k, 3*2^k or 5*2^k in log2"
                                 if rounding == "smooth" else
                                 "side log-uniform, rounded to the nearest power of two in log2"),
                    "steps": "U{500..5000}", "sigma": "U[0.05,0.2]", "n0": "U[0.02,0.08]"})
    finally:
        svc.close()
    return rep


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the 2-D 256^2 and 3-D 256^3 FP64 configuration runs")
    ap.add_argument("--serve", type=float, default=30.0,
                    help="seconds of Poisson arrivals for the serving leg (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-3d", type=float, default=0.5, help="fraction of 3-D jobs")
    ap.add_argument("--serve-class-cells", type=int, default=1 << 25,
                    help="slot policy: cells per size-class launch")
    ap.add_argument("--serve-lookahead", type=int, default=1,
                    help="blocks handed at once to a size class no queued job waits for")
    ap.add_argument("--serve-lanes", default="class",
                    help="'class' (a lane, with its own stream and block cadence, per size class), "
                         "'one' (one lane for all), or groups of classes such as "
                         "'2d_128,2d_256;2d_512,2d_1024;3d_64,3d_128;3d_256'")
    ap.add_argument("--serve-rounding", choices=["smooth", "pow2"], default="smooth",
                    help="job sides rounded to 2^k, 3*2^k, 5*2^k (smooth) or to powers of two")
    ap.add_argument("--only", choices=["batch", "serve", "configs"],
                    help="run only that leg and print its JSON")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="run --steps unsampled steps and one field read-back (its D2D copy of "
                         "the packed spectrum is the known-size copy for counter calibration)")
    ap.add_argument("--variants", default="",
                    help="pass variants 'x,y,z' of the 512^3 FP32 run (and the traffic probe) "
                         "to pin instead of autotuning")
    args = ap.parse_args()
    if args.traffic_probe:
        import xpfc_cuda_amd as X
        p = fcc_params(args.n, 1, t_sample=0)
        with X.Context(p) as ctx:
            if args.variants:
                ctx.set_variants(*[int(v) for v in args.variants.split(",")])
            ctx.initialize_noise(0.01, 2)
            ctx.step(args.steps)
            ctx.get_field()
        print(json.dumps({"traffic_probe": True, "n": args.n, "steps": args.steps,
                          "copy_bytes": spectrum_bytes(args.n, 1),
                          "xstep_algorithmic_bytes": XSTEP_SPECTRA * spectrum_bytes(args.n, 1)}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the workload would shard as x-slabs with two all-to-all transposes per step)",
              file=sys.stderr)
        sys.exit(2)
    if args.only == "configs":
        print(json.dumps({"configs": config_runs()}), flush=True)
        return
    if args.only == "batch":
        print(json.dumps({"batched": batched_run()}), flush=True)
        return
    if args.only == "serve":
        print(json.dumps({"serving": serving_run(args.serve, args.serve_rate, args.serve_3d,
                                                  class_cells=args.serve_class_cells,
                                                  lookahead=args.serve_lookahead,
                                                  lanes=args.serve_lanes,
                                                  rounding=args.serve_rounding)}), flush=True)
        return
    n = args.n
    pin = [int(v) for v in args.variants.split(",")] if args.variants else None
    ms, wall, passes, samples, variants, io = gpu_run(n, 1, args.steps, args.warmup, pin=pin)
    ms_per_step = ms / args.steps
    value = 1000.0 / ms_per_step
    S = spectrum_bytes(n, 1)
    # dominant kernel: the axis-0 fused pass; 4 S algorithmic bytes per launch
    # (the model's figure for the pass), 3 S of HBM traffic in this state form
    x_ms = passes[0]
    achieved = XSTEP_MODEL_SPECTRA * S / (x_ms * 1e-3) / 1e9
    x_hbm = XSTEP_SPECTRA * S / (x_ms * 1e-3) / 1e9
    step_gbs = STEP_SPECTRA_3D * S / (passes[4] * 1e-3) / 1e9
    step_model_gbs = STEP_MODEL_SPECTRA_3D * S / (passes[4] * 1e-3) / 1e9
    out = {
        "metric": "XPFC steps/s at 512^3 (FP32-mixed)",
        "value": round(value, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (fp64 denominator, k=0 mode, reductions)",
        "data": "synthetic: n0=0.05+0.01*N(0,1), seed 2",
        "config": {"workload": "3-D fcc two-peak XPFC %d^3, dx=0.125, dt=0.5, sigma=0.1, "
                               "FP32-mixed, FE/RMSE every 100 steps" % n,
                   "grid": [n, n, n], "parallelism": "single GPU"},
        "cell_updates_per_s": value * n ** 3,
        "step_hbm_gbs": round(step_gbs, 1),
        "step_model_gbs": round(step_model_gbs, 1),
        "pass_ms": {"x_fused": passes[0], "y_inv": passes[1], "z_real": passes[2],
                    "y_fwd": passes[3], "step": passes[4]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None,
                     "kernel": "k_xstep (axis-0 FFTs of the nonlinear term and the state, update, inverse FFT)",
                     "algorithmic_bytes": XSTEP_MODEL_SPECTRA * S,
                     "hbm_bytes": XSTEP_SPECTRA * S, "hbm_gbs": round(x_hbm, 1),
                     "hbm_frac": round(x_hbm / HBM_PEAK_GBS, 4),
                     "note": "algorithmic_bytes = the pass's 4 S of the compulsory-traffic model; "
                             "with the state held along axis 0 the kernel moves 3 S (hbm_bytes, "
                             "measured: traffic), so achieved is the model's rate and hbm_gbs the HBM rate"},
        "samples": [s.line().strip() for s in samples[:4]],
    }
    # the practical ceiling: a plain streaming copy of one spectrum (read + write)
    copy = copy_ceiling(S)
    out["roofline"]["copy_gbs"] = round(copy, 1)
    out["roofline"]["frac_of_copy"] = round(x_hbm / copy, 4)
    out["step_frac_of_copy"] = round(step_gbs / copy, 4)
    if not args.no_fp64:
        fsteps = max(10, args.steps // 4)
        ms64, _, passes64, _, _, _ = gpu_run(n, 0, fsteps, min(args.warmup, 50))
        v64 = 1000.0 / (ms64 / fsteps)
        out["fp64_steps_per_s"] = round(v64, 3)
        out["fp32_over_fp64"] = round(value / v64, 3)
        out["fp64_step_hbm_gbs"] = round(STEP_SPECTRA_3D * spectrum_bytes(n, 0) / (passes64[4] * 1e-3) / 1e9, 1)
        out["accuracy_vs_fp64"] = accuracy_run(n)
    out["variants"] = {"axis0_tile": variants[0], "axis1_tile": variants[1] & ~1,
                       "axis1_xcd_pairs": bool(variants[1] & 1),
                       "real_space_block": variants[2] & ~1,
                       "real_space_free_regs": bool(variants[2] & 1)}
    if io:
        # a 1000-step job that takes its field from host memory and returns it
        io_ms = io["get_field_ms"] + io["set_field_ms"]
        io["steps_per_s_1000_step_job_incl_host_io"] = round(1000.0 / (1000 * ms_per_step + io_ms) * 1e3, 2)
        out["host_io"] = io
    tp = os.path.join(ROOT, "profiles", "traffic_512cube_fp32.json")
    if n == 512 and os.path.exists(tp):
        # HBM bytes per k_xstep launch of the variant in use, from the committed
        # FETCH_SIZE / WRITE_SIZE runs
        with open(tp) as f:
            tr = json.load(f)["k_xstep"].get(str(variants[0]))
        if tr:
            out["roofline"]["traffic"] = tr["traffic_bytes"]
    if not args.no_configs:
        out["configs"] = config_runs()
    if not args.no_batch:
        out["batched"] = batched_run()
    if args.serve > 0:
        out["serving"] = serving_run(args.serve, args.serve_rate, args.serve_3d,
                                     class_cells=args.serve_class_cells,
                                     lookahead=args.serve_lookahead, lanes=args.serve_lanes,
                                     rounding=args.serve_rounding)
    if not args.no_cpu:
        out["cpu_baseline"], psi0, ref, nref, workers = cpu_baseline(n)
        out["accuracy_vs_oracle"] = accuracy_vs_oracle(n, psi0, ref, nref, workers)
        del psi0, ref
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
