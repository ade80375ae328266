"""Golden fixtures (tests/golden/*.npz, written by gen_golden.py from the
oracle): the oracle still reproduces them (CPU), and the HIP path matches them
through the C ABI (GPU)."""
import glob
import math
import os

import numpy as np
import pytest

from oracle import xpfc_oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(HERE, "*.npz"))
               if "philox" not in p)


def load(name):
    return np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False)


def oracle_params(z):
    n = tuple(int(v) for v in z["n"])
    return O.Params(n=n, L=tuple(m * 0.125 for m in n), sigma=float(z["sigma"]), psi0=0.05,
                    peaks=[O.Peak(*map(float, pk)) for pk in z["peaks"]],
                    dcf_mode=str(z["mode"]), t_sample=int(z["t_sample"]))


def test_fixtures_present():
    assert len(CASES) >= 3


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_fixture(name):
    z = load(name)
    p = oracle_params(z)
    psi0 = O.noise_ic(p, float(z["amp"]), int(z["seed"]))
    assert np.array_equal(psi0, z["psi0"])
    out, s = O.run(psi0, p, int(z["steps"]), sample=True)
    assert np.max(np.abs(out - z["psi_final"])) < 1e-15
    fe = np.array([[x.step, x.value, x.f_mean, x.mean] for x in s if x.kind == "FE"])
    assert np.allclose(fe, z["fe"], rtol=1e-13, atol=0)


def test_philox_normals_fixture():
    z = np.load(os.path.join(HERE, "philox_normals.npz"), allow_pickle=False)
    assert np.array_equal(O.philox_normals(64, 0), z["seed0"])
    assert np.array_equal(O.philox_normals(64, (1 << 40) + 3), z["seed_big"])


def test_library_randn_matches_fixture():
    import xpfc_cuda_amd as X
    z = np.load(os.path.join(HERE, "philox_normals.npz"), allow_pickle=False)
    assert np.max(np.abs(X.randn(0, 64) - z["seed0"])) < 4e-15
    assert np.max(np.abs(X.randn((1 << 40) + 3, 64) - z["seed_big"])) < 4e-15


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("name", CASES)
def test_gpu_matches_fixture(name, prec):
    import xpfc_cuda_amd as X
    z = load(name)
    n = tuple(int(v) for v in z["n"])
    print("golden %s prec=%d" % (name, prec), flush=True)
    lp = X.Params(n=n, L=tuple(m * 0.125 for m in n), sigma=float(z["sigma"]), psi0=0.05,
                  peaks=tuple(tuple(map(float, pk)) for pk in z["peaks"]),
                  dcf_mode=X.XPFC_DCF_MAX if str(z["mode"]) == "max" else X.XPFC_DCF_SUM,
                  precision=prec, t_sample=int(z["t_sample"]))
    with X.Context(lp) as ctx:
        ctx.initialize_noise(float(z["amp"]), int(z["seed"]))
        psi0 = ctx.get_field()
        assert np.max(np.abs(psi0 - z["psi0"])) <= (1e-15 if prec == 0 else 5e-8)
        ctx.step(int(z["steps"]))
        got = ctx.get_field()
        gs = ctx.samples()
    tol = 1e-11 if prec == 0 else 1e-5
    assert np.max(np.abs(got - z["psi_final"])) <= tol
    gfe = {s.step: s for s in gs if s.has_fe}
    for st, F, fm, mean in z["fe"]:
        g = gfe[int(st)]
        assert abs(g.F_ref - F) <= (1e-11 if prec == 0 else 1e-7) * abs(F)
        assert abs(g.mean - mean) <= (1e-15 if prec == 0 else 1e-9)
