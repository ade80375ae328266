"""The 512^3 benchmark workload after 1000 steps against the FP64 oracle.

The check the north-star target names (max |dn| <= 1e-5, relative free-energy
error <= 1e-7 after 1000 steps at 512^3) is too long for the test suite: the
oracle takes about 25 CPU-minutes with 16 threads.  It runs in two halves:

    python tests/accuracy_512cube.py gpu DIR          # on the GPU: writes DIR/gpu.npz
    python tests/accuracy_512cube.py oracle DIR [W]   # on any host: runs the oracle
                                                      # (W threads), writes DIR/oracle.npz,
                                                      # compares, prints JSON
    python tests/accuracy_512cube.py compare DIR      # compares DIR/gpu.npz with an
                                                      # existing DIR/oracle.npz

The GPU half runs the workload (fcc two-peak kernel, sigma 0.1, n0 0.05,
noise 0.01 from the seed-2 Philox draws, FE/RMSE every 100 steps) for 1000
steps in FP32-mixed and in FP64 and keeps the free-energy / RMSE / mean
samples and every 8th point of the field along each axis (64^3 values), so the
file stays small.  The oracle half starts from the same draws (noise_ic, the
host form of the device generator), runs the reference's loop with the same
sampling schedule, keeps the same subsample of its field and its samples, and
reports max |dn| over the kept points, the relative free-energy difference at
every sample and the RMSE samples.  Not collected by
pytest (no test_ prefix); the result is summarised under profiles/.
"""
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

N, STEPS, T_SAMPLE, SEED, AMP = 512, 1000, 100, 2, 0.01
SUB = 8  # keep every SUB-th point along each axis


def gpu(outdir):
    import xpfc_cuda_amd as X
    k1 = 2.0 * math.pi
    L = N * 0.125
    out = {}
    for prec, key in ((X.XPFC_FP32_MIXED, "fp32"), (X.XPFC_FP64, "fp64")):
        p = X.Params(n=(N, N, N), L=(L, L, L), dt=0.5, Mo=1.0, eps=0.0, a=1.0, b=1.0 / 3.0,
                     sigma=0.1, psi0=0.05, peaks=((k1, 1.0, 1.0), (k1 * 2.0 / math.sqrt(3.0), 1.0, 1.0)),
                     dcf_mode=X.XPFC_DCF_MAX, precision=prec, t_sample=T_SAMPLE)
        with X.Context(p) as ctx:
            ctx.initialize_noise(AMP, SEED)
            out[key + "_init"] = ctx.get_field()[::SUB, ::SUB, ::SUB].copy()
            t0 = time.perf_counter()
            ctx.step(STEPS)
            ctx.sync()
            print("%s: %d steps in %.2f s" % (key, STEPS, time.perf_counter() - t0), flush=True)
            f = ctx.get_field()
            out[key + "_field"] = f[::SUB, ::SUB, ::SUB].copy()
            out[key + "_mean"] = np.array(float(np.mean(f)))
            del f
            smp = ctx.samples()
            F_end, _ = ctx.fe_calc()
        out[key + "_fe"] = np.array([(s.step, s.F_ref) for s in smp if s.has_fe] + [(STEPS, F_end)])
        out[key + "_sse"] = np.array([(s.step, s.rmse) for s in smp if s.has_sse])
    os.makedirs(outdir, exist_ok=True)
    np.savez_compressed(os.path.join(outdir, "gpu.npz"), **out)
    print("wrote", os.path.join(outdir, "gpu.npz"), flush=True)


def oracle(outdir, workers):
    from oracle import xpfc_oracle as O
    k1 = 2.0 * math.pi
    L = N * 0.125
    op = O.Params(n=(N, N, N), L=(L, L, L), sigma=0.1, psi0=0.05,
                  peaks=[O.Peak(k1, 1.0, 1.0), O.Peak(k1 * 2.0 / math.sqrt(3.0), 1.0, 1.0)],
                  dcf_mode="max", t_sample=T_SAMPLE)
    psi0 = O.noise_ic(op, AMP, SEED)
    init = psi0[::SUB, ::SUB, ::SUB].copy()
    t0 = time.perf_counter()
    ref, rs = O.run(psi0, op, STEPS, sample=True, workers=workers)
    dt = time.perf_counter() - t0
    del psi0
    np.savez_compressed(os.path.join(outdir, "oracle.npz"), init=init, field=ref[::SUB, ::SUB, ::SUB].copy(),
                        mean=np.array(float(np.mean(ref))),
                        fe=np.array([(s.step, s.value) for s in rs if s.kind == "FE"]),
                        sse=np.array([(s.step, s.value) for s in rs if s.kind == "SSE"]),
                        seconds=np.array(dt), threads=np.array(workers))
    compare(outdir)


def compare(outdir):
    g = np.load(os.path.join(outdir, "gpu.npz"))
    o = np.load(os.path.join(outdir, "oracle.npz"))
    init_diff = {k: float(np.max(np.abs(g[k + "_init"] - o["init"]))) for k in ("fp32", "fp64")}
    ofe = {int(s): v for s, v in o["fe"]}
    osse = {int(s): v for s, v in o["sse"]}
    sub = o["field"]
    res = {"workload": "3-D fcc two-peak 512^3, sigma 0.1, n0 0.05, noise 0.01 seed 2, %d steps" % STEPS,
           "oracle_seconds": round(float(o["seconds"]), 1), "oracle_threads": int(o["threads"]),
           "points_compared": int(sub.size), "initial_field_max_abs_diff": init_diff,
           "oracle_mean": float(o["mean"])}
    for k in ("fp32", "fp64"):
        fe = {int(s): v for s, v in g[k + "_fe"]}
        sse = {int(s): v for s, v in g[k + "_sse"]}
        rel = {s: abs(fe[s] - ofe[s]) / abs(ofe[s]) for s in sorted(ofe) if s in fe}
        srel = {s: abs(sse[s] - osse[s]) / abs(osse[s]) for s in sorted(osse) if s in sse}
        res[k] = {"max_abs_dn_subsampled": float(np.max(np.abs(g[k + "_field"] - sub))),
                  "fe_rel_at_end": rel.get(STEPS), "fe_rel_max_over_samples": max(rel.values()),
                  "rmse_rel_max_over_samples": max(srel.values()) if srel else None,
                  "mean_minus_oracle": float(g[k + "_mean"]) - res["oracle_mean"],
                  "F_end": fe[STEPS], "F_end_oracle": ofe[STEPS]}
    print(json.dumps(res, indent=1), flush=True)
    with open(os.path.join(outdir, "compare.json"), "w") as f:
        json.dump(res, f, indent=1)


if __name__ == "__main__":
    if len(sys.argv) < 3 or sys.argv[1] not in ("gpu", "oracle", "compare"):
        sys.exit(__doc__)
    if sys.argv[1] == "gpu":
        gpu(sys.argv[2])
    elif sys.argv[1] == "compare":
        compare(sys.argv[2])
    else:
        oracle(sys.argv[2], int(sys.argv[3]) if len(sys.argv) > 3 else (os.cpu_count() or 1))
