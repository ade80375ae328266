"""The 512^3 workload at full size, through properties that do not need the
oracle (which is far too slow at this size): the closed-form evolution of a
linear two-mode field, and bitwise agreement of different pass variants and of
graph replays with eager launches on the nonlinear workload, with the mass
mode unchanged."""
import math

import numpy as np
import pytest

from xpfc_cuda_amd import Context, Params, XPFC_DCF_MAX, XPFC_FP32_MIXED, XPFC_FP64

pytestmark = pytest.mark.gpu

N = 512
L = 64.0
TWO_PI = 2.0 * math.pi
PEAKS = ((TWO_PI, 1.0, 1.0), (TWO_PI * 2.0 / math.sqrt(3.0), 1.0, 1.0))  # fcc


def _flush(msg):
    print(msg, flush=True)


def _params(prec, sigma, a, b, t_sample=0):
    return Params(n=(N, N, N), L=(L, L, L), dt=0.5, Mo=1.0, eps=0.0, a=a, b=b, sigma=sigma,
                  psi0=0.05, peaks=PEAKS, dcf_mode=XPFC_DCF_MAX, precision=prec,
                  t_sample=t_sample)


def _plane_wave(idx, amp, out):
    """out += amp * cos(k . x) with k = 2 pi / L * idx on the 512^3 grid."""
    x = np.arange(N) * (L / N)
    kx, ky, kz = (TWO_PI / L * i for i in idx)
    ph = (kx * x)[:, None, None] + (ky * x)[None, :, None] + (kz * x)[None, None, :]
    np.cos(ph, out=ph)
    ph *= amp
    out += ph


@pytest.mark.parametrize("prec,tol", [(XPFC_FP64, 1e-13), (XPFC_FP32_MIXED, 1e-7)])
def test_linear_two_modes_closed_form_512cube(prec, tol):
    """a = b = 0 (no nonlinearity), sigma = 0: each Fourier mode is multiplied
    by 1 / (1 + dt M k^2 (1 - C2(k))) per step, C2(k) = max_i exp(-(|k| - k_i)^2
    / (2 alpha_i^2)).  One mode far below the peaks, one close to the first."""
    _flush("512^3 linear closed form prec=%d" % prec)
    p = _params(prec, sigma=0.0, a=0.0, b=0.0)
    modes = [((3, 2, 1), 0.01), ((48, 32, 24), 0.008)]
    psi = np.full((N, N, N), 0.05)
    for idx, amp in modes:
        _plane_wave(idx, amp, psi)
    mean0 = float(np.mean(psi))  # 0.05 up to the rounding of the cosines
    nsteps = 10
    with Context(p) as c:
        c.set_field(psi)
        c.step(nsteps)
        got = c.get_field()
        mean = float(np.mean(got))   # of the field read back, not the carried scalar
    expect = np.full((N, N, N), 0.05)
    for idx, amp in modes:
        k = TWO_PI / L * math.sqrt(sum(i * i for i in idx))
        c2 = max(math.exp(-(k - kp) ** 2 / (2.0 * al * al)) for kp, al, _ in PEAKS)
        g = 1.0 / (1.0 + 0.5 * k * k * (1.0 - c2))
        assert 0.1 < g ** nsteps < 0.9  # both modes measurably damped
        _plane_wave(idx, amp * g ** nsteps, expect)
    assert np.max(np.abs(got - expect)) <= tol
    # the mass mode is carried unchanged: the field read back (mean + fluctuation,
    # the fluctuation in the working precision) has the initial mean to the
    # rounding of summing 1.3e8 cells (FP64) / of the FP32 fluctuations
    assert mean == pytest.approx(mean0, abs=1e-13 if prec == XPFC_FP64 else 1e-10)


def test_variants_and_graphs_bitwise_512cube():
    """The nonlinear fcc workload (FP32-mixed): two different variant sets, one
    run from graph replays and one launched eagerly, give the same field bit
    for bit after 12 steps; the mean stays at n0."""
    _flush("512^3 variants / graphs bitwise")
    p = _params(XPFC_FP32_MIXED, sigma=0.1, a=1.0, b=1.0 / 3.0, t_sample=4)
    out = []
    for v, graphs in [((8, 17, 256), True), ((16, 16, 129), False)]:
        with Context(p) as c:
            c.set_variants(*v)
            c.set_graphs(graphs)
            c.initialize_noise(0.01, 2)
            c.step(12)
            f = c.get_field()
            out.append(f)
            # the real-space mean the kernels sum at each sample, and the
            # field's own mean, stay at n0 (the noise is drawn with its mean
            # removed; FP32 fluctuations summed in FP64)
            ms = [s.mean for s in c.samples() if s.has_fe]
            assert len(ms) >= 3
            for m in ms:
                assert m == pytest.approx(0.05, abs=1e-10)
            assert float(np.mean(f)) == pytest.approx(0.05, abs=1e-10)
            F, _ = c.fe_calc()
            assert math.isfinite(F)
    assert np.array_equal(out[0], out[1])


def test_fp32_mixed_vs_fp64_1000_steps_512cube():
    """The 512^3 fcc workload (sigma = 0.1, n0 = 0.05 + 0.01 N(0,1)) after
    1000 steps in FP32-mixed and in FP64 from the same initial field: max |dn|
    <= 1e-5 and relative free-energy difference <= 1e-7.  (The FP64 path is
    held to the FP64 oracle at smaller sizes; the oracle itself would take
    about 25 CPU-minutes here.)"""
    _flush("512^3 FP32-mixed vs FP64, 1000 steps")
    fields, fe = [], []
    for prec in (XPFC_FP32_MIXED, XPFC_FP64):
        with Context(_params(prec, sigma=0.1, a=1.0, b=1.0 / 3.0)) as c:
            c.initialize_noise(0.01, 2)
            c.step(1000)
            fields.append(c.get_field())
            fe.append(c.fe_calc()[0])
    dn = float(np.max(np.abs(fields[0] - fields[1])))
    print("max|dn| = %.3e, FE rel = %.3e" % (dn, abs(fe[0] - fe[1]) / abs(fe[1])), flush=True)
    assert dn <= 1e-5
    assert abs(fe[0] - fe[1]) <= 1e-7 * abs(fe[1])
