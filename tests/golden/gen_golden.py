"""Write the golden fixtures under tests/golden/ from the FP64 oracle.

The reference ships no fixtures (no tests, no sample inputs), so these are the
oracle's outputs on seeded inputs; the oracle itself is pinned by
tests/test_oracle.py.  Run: python tests/golden/gen_golden.py
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import xpfc_oracle as O  # noqa: E402

TWO_PI = 2.0 * math.pi

CASES = {
    # crystallising 2-D square lattice (max envelope, ρβ = 1), growth phase
    "sq2d_32x32_s007": dict(n=(32, 32), sigma=0.07, steps=60, t_sample=10, seed=0, amp=0.01,
                            peaks=[(TWO_PI, 1.0, 1.0), (TWO_PI * math.sqrt(2), 1.0, 1.0)],
                            mode="max"),
    # the reference's own DCF (sum, ρβ = 4, 4/√2), stable liquid at σ = 0.2
    "ref2d_64x32_sum": dict(n=(64, 32), sigma=0.2, steps=50, t_sample=25, seed=5, amp=0.02,
                            peaks=[(pk.k, pk.alpha, pk.rho_beta) for pk in O.reference_peaks()],
                            mode="sum"),
    # 3-D bcc single peak, growth phase
    "bcc3d_16x16x32": dict(n=(16, 16, 32), sigma=0.07, steps=30, t_sample=10, seed=1, amp=0.01,
                           peaks=[(TWO_PI, 1.0, 1.0)], mode="max"),
}


def oracle_params(c):
    n = c["n"]
    return O.Params(n=n, L=tuple(m * 0.125 for m in n), sigma=c["sigma"], psi0=0.05,
                    peaks=[O.Peak(*pk) for pk in c["peaks"]], dcf_mode=c["mode"],
                    t_sample=c["t_sample"])


def main():
    for name, c in CASES.items():
        p = oracle_params(c)
        psi0 = O.noise_ic(p, c["amp"], c["seed"])
        out, samples = O.run(psi0, p, c["steps"], sample=True)
        fe = np.array([[s.step, s.value, s.f_mean, s.mean] for s in samples if s.kind == "FE"])
        sse = np.array([[s.step, s.value] for s in samples if s.kind == "SSE"])
        np.savez_compressed(os.path.join(HERE, name + ".npz"), psi0=psi0, psi_final=out,
                            fe=fe, sse=sse, n=np.array(c["n"]), sigma=c["sigma"],
                            steps=c["steps"], t_sample=c["t_sample"], seed=c["seed"],
                            amp=c["amp"], peaks=np.array(c["peaks"]),
                            mode=np.array(c["mode"]))
        print(name, "max|dpsi| =", float(np.max(np.abs(out - psi0))))
    z = {"seed0": O.philox_normals(64, 0), "seed_big": O.philox_normals(64, (1 << 40) + 3)}
    np.savez_compressed(os.path.join(HERE, "philox_normals.npz"), **z)


if __name__ == "__main__":
    main()
