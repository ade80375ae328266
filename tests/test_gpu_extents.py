"""GPU parity at grid extents beyond powers of two: sides of 3*2^k and 5*2^k
points (radix-3/5/6/10/12/20 butterflies) and lines of up to 2048 points
(FP64 2048-point lines use 2-column tiles), through the C ABI against numpy's
FFT and the FP64 oracle.  The reference takes any Mx, My (read_input.cu:40-41)
and lets cuFFT plan them (main.cu:106-113).

Tolerances as test_gpu_parity.py: transforms 1e-12 (FP64) / 2e-6 (FP32) of
the spectrum's scale; steps max|Δn| <= 1e-10 (FP64) / 1e-5 (FP32-mixed), FE
samples 1e-11 / 1e-7 relative.
"""
import numpy as np
import pytest

from oracle import xpfc_oracle as O
from xpfc_cuda_amd import Context, XPFCError, XPFC_FP32_MIXED, XPFC_FP64

from cases import growth_case, make

pytestmark = pytest.mark.gpu

PRECS = [XPFC_FP64, XPFC_FP32_MIXED]
TOL_FFT = {XPFC_FP64: 1e-12, XPFC_FP32_MIXED: 2e-6}

# every new line length appears along every axis at least once: 3*2^k and
# 5*2^k (12 .. 1536, 20 .. 1280) as strided lines and as half lines, 2048 along
# axis 0 and axis 1, 1024-point half lines (n2 = 2048)
SHAPES = [(12, 24), (20, 40), (96, 96), (160, 80), (48, 1536), (1536, 12), (1280, 20), (20, 1280),
          (2048, 16), (16, 2048), (384, 640), (768, 320), (12, 20, 24), (96, 96, 96), (80, 48, 96),
          (8, 1280, 16), (16, 8, 1280), (20, 12, 40), (8, 2048, 16), (192, 160, 24), (40, 768, 12)]


def _flush(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("prec", PRECS)
def test_fft_r2c_c2r_vs_numpy(shape, prec):
    _flush("fft %s prec=%d" % (shape, prec))
    _, lp = make(shape, prec=prec)
    rng = np.random.default_rng(11)
    x = rng.standard_normal(shape)
    with Context(lp) as ctx:
        X = ctx.fft_r2c(x)
        ref = np.fft.rfftn(x)
        scale = np.max(np.abs(ref))
        assert np.max(np.abs(X - ref)) <= TOL_FFT[prec] * scale * 4
        back = ctx.fft_c2r(ref)
        assert np.max(np.abs(back / x.size - x)) <= TOL_FFT[prec] * np.max(np.abs(x)) * 8


@pytest.mark.parametrize("shape", [(12, 12), (20, 20), (12, 20, 24), (20, 12, 12)])
@pytest.mark.parametrize("prec", PRECS)
def test_smallest_new_grids_vs_oracle(shape, prec):
    """12- and 20-point lines (one thread per line, one radix-12/20 stage)
    and 6- and 10-point half lines, after 10 steps with a sample every step."""
    _flush("small %s prec=%d" % (shape, prec))
    op, lp = growth_case(shape, prec=prec, t_sample=1)
    psi0 = O.noise_ic(op, 0.02, 4)
    ref, rs = O.run(psi0, op, 10, sample=True)
    with Context(lp) as ctx:
        ctx.set_field(psi0)
        ctx.step(10)
        got = ctx.get_field()
        gs = ctx.samples()
    assert np.max(np.abs(got - ref)) <= (1e-12 if prec == XPFC_FP64 else 1e-6)
    rfe = {s.step: s.value for s in rs if s.kind == "FE"}
    gfe = {s.step: s.F_ref for s in gs if s.has_fe}
    assert sorted(gfe) == sorted(rfe) and rfe
    for st, r in rfe.items():
        assert abs(gfe[st] - r) <= (1e-11 if prec == XPFC_FP64 else 1e-7) * abs(r), (st, gfe[st], r)


# (96, 96, 96): the cubic table path with 3*2^k lines on every axis;
# (2048, 2048): the longest 2-D lines (FP64: 2-column axis-0 tiles);
# (160, 80), (80, 96, 64), (192, 40, 24): 5*2^k and mixed factors, non-square boxes
@pytest.mark.parametrize("shape,nsteps", [((96, 96, 96), 20), ((2048, 2048), 10), ((160, 80), 60),
                                          ((80, 96, 64), 20), ((192, 40, 24), 20), ((1536, 48), 20),
                                          ((20, 1280), 20)])
@pytest.mark.parametrize("prec", PRECS)
def test_steps_vs_oracle(shape, nsteps, prec):
    _flush("steps %s n=%d prec=%d" % (shape, nsteps, prec))
    op, lp = growth_case(shape, prec=prec, t_sample=10)
    psi0 = O.noise_ic(op, 0.01, 0)
    ref, rs = O.run(psi0, op, nsteps, sample=True)
    with Context(lp) as ctx:
        ctx.set_field(psi0)
        ctx.step(nsteps)
        got = ctx.get_field()
        gs = ctx.samples()
    assert np.max(np.abs(ref - psi0)) > 1e-4  # the field evolved
    err = np.max(np.abs(got - ref))
    assert err <= (1e-10 if prec == XPFC_FP64 else 1e-5), err
    rfe = {s.step: s for s in rs if s.kind == "FE"}
    gfe = {s.step: s for s in gs if s.has_fe}
    rsse = {s.step: s.value for s in rs if s.kind == "SSE"}
    gsse = {s.step: s.rmse for s in gs if s.has_sse}
    assert set(rfe) == set(gfe) and sorted(rsse) == sorted(gsse) and rfe
    ftol = 1e-11 if prec == XPFC_FP64 else 1e-7
    for st, r in rfe.items():
        assert abs(gfe[st].F_ref - r.value) <= ftol * abs(r.value), (st, gfe[st].F_ref, r.value)
        assert abs(gfe[st].mean - r.mean) <= (1e-14 if prec == XPFC_FP64 else 1e-9)
    for st, r in rsse.items():
        assert abs(gsse[st] - r) <= (1e-9 if prec == XPFC_FP64 else 1e-4) * abs(r) + 1e-300, (st, gsse[st], r)


@pytest.mark.parametrize("shape", [(96, 96, 96), (160, 320)])
@pytest.mark.parametrize("prec", PRECS)
def test_variants_bitwise_identical(shape, prec):
    """Every tile width and real-space block gives the same field bit for bit
    (they regroup lines, never change a line's arithmetic)."""
    _flush("variants %s prec=%d" % (shape, prec))
    op, lp = growth_case(shape, prec=prec, t_sample=5)
    psi0 = O.noise_ic(op, 0.01, 2)
    fields = []
    with Context(lp) as ctx:
        for tc in (16, 8, 4, 2):
            for ntz in (64, 256, 512):
                try:
                    ctx.set_variants(tc, tc if len(shape) == 3 else 0, ntz)
                except XPFCError:
                    continue
                ctx.set_field(psi0)
                ctx.step(12)
                fields.append(((tc, ntz), ctx.get_field()))
    assert len(fields) >= 3
    for v, f in fields[1:]:
        assert np.array_equal(f, fields[0][1]), v


# steps at which the oracle's amplitude max|n - n0| has just passed 0.1
# (fcc two-peak kernel, sigma 0.07, noise 0.01 seed 2)
CRYST_CASES = [((96, 96, 96), 120), ((80, 80, 80), 110)]
# relative to the amplitude, as for the power-of-two crystallising cases
CRYST_TOL = {XPFC_FP64: 1e-11, XPFC_FP32_MIXED: 3e-4}


@pytest.mark.parametrize("n,nsteps", CRYST_CASES)
@pytest.mark.parametrize("prec", PRECS)
def test_crystallising_radix3_radix5_vs_oracle(n, nsteps, prec):
    """The fcc two-peak physics of the 512^3 workload made crystallising
    (sigma 0.07) on 96^3 (radix 3) and 80^3 (radix 5) grids, grown from noise
    until the amplitude passes 0.1 (the strongly nonlinear regime), against
    the oracle: field and FE / RMSE samples."""
    import math
    _flush("crystallising %s steps=%d prec=%d" % (n, nsteps, prec))
    k1 = 2.0 * math.pi
    op, lp = make(n, sigma=0.07, psi0=0.05, peaks=[(k1, 1.0, 1.0), (k1 * 2.0 / math.sqrt(3.0), 1.0, 1.0)],
                  mode="max", prec=prec, t_sample=20)
    psi0 = O.noise_ic(op, 0.01, 2)
    ref, rs = O.run(psi0, op, nsteps, sample=True)
    amp = float(np.max(np.abs(ref - 0.05)))
    assert 0.1 < amp < 0.5
    with Context(lp) as ctx:
        ctx.set_field(psi0)
        ctx.step(nsteps)
        got = ctx.get_field()
        gs = ctx.samples()
    err = float(np.max(np.abs(got - ref)))
    _flush("crystallising %s prec=%d amp=%.3g max|dn|=%.3e rel=%.3e" % (n, prec, amp, err, err / amp))
    assert err <= CRYST_TOL[prec] * amp
    fp64 = prec == XPFC_FP64
    rfe = {s.step: s.value for s in rs if s.kind == "FE" and s.step < nsteps}
    gfe = {s.step: s.F_ref for s in gs if s.has_fe and s.step < nsteps}
    assert set(rfe) == set(gfe) and rfe
    for st, r in rfe.items():
        assert abs(gfe[st] - r) <= (1e-10 if fp64 else 1e-5) * abs(r), (st, gfe[st], r)


def test_largest_planes_both_state_forms(monkeypatch):
    """2048 x 2048 x 1280 FP32 (64 GB of spectra): 2048-point axis-0 lines over
    planes of 1.3M elements, where a line's row offsets pass 2^31 elements.
    From device noise, three steps in each state form: finite, mass exact,
    the free energy falling step by step (a stable liquid), and the two forms
    agreeing to FP32 rounding.  (No host copy of the 5.4G-point field.)"""
    import math
    n = (2048, 2048, 1280)
    k1 = 2 * math.pi
    out = {}
    for form in ("0", "1"):
        monkeypatch.setenv("XPFC_STATE_FORM", form)
        monkeypatch.setenv("XPFC_AUTOTUNE", "0")
        _flush("largest planes %s form %s" % (n, form))
        _, p = make(n, sigma=0.1, psi0=0.05, peaks=[(k1, 1.0, 1.0), (k1 * 2.0 / math.sqrt(3.0), 1.0, 1.0)],
                    prec=XPFC_FP32_MIXED, t_sample=0)
        Fs = []
        with Context(p) as ctx:
            ctx.initialize_noise(0.01, 3)
            Fs.append(ctx.fe_calc()[0])
            for _ in range(3):
                ctx.step(1)
                Fs.append(ctx.fe_calc()[0])
                assert abs(ctx.avg_calc() - 0.05) <= 1e-15
        _flush("largest planes form %s F %s" % (form, Fs))
        assert all(math.isfinite(F) for F in Fs)
        assert all(b < a for a, b in zip(Fs, Fs[1:]))
        out[form] = Fs
    for a, b in zip(out["0"], out["1"]):
        assert abs(a - b) <= 1e-6 * abs(a)


def _hip_mem_info():
    """(free, total) device bytes from the HIP runtime the library runs on."""
    import ctypes as ct
    path = None
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64.so" in line:
                path = line.split()[-1]
                break
    assert path, "HIP runtime not loaded"
    hip = ct.CDLL(path)
    free, total = ct.c_size_t(), ct.c_size_t()
    assert hip.hipMemGetInfo(ct.byref(free), ct.byref(total)) == 0
    return free.value, total.value


def test_autotune_in_place_when_device_memory_is_short(monkeypatch):
    """The autotuner times the pass variants in two scratch spectra; when free
    device memory cannot hold them beside the context (the largest grids,
    whose three buffers fill most of the GPU) it times them in place in the
    context's diagnostics buffer.  Here the memory is filled from outside
    before the first step: tuning succeeds, later calls see no stale
    allocation error, and the field equals an untuned run's bit for bit (the
    variants only regroup lines)."""
    from xpfc_cuda_amd.api import DeviceBuffer
    shape = (160, 96, 128)   # a grid no other test tunes (tuning results are cached per grid)
    _flush("autotune in place %s" % (shape,))
    _, p = make(shape, prec=XPFC_FP32_MIXED, sigma=0.09, psi0=0.05, t_sample=10)
    psi0 = 0.05 + 0.02 * np.random.default_rng(5).standard_normal(shape)
    monkeypatch.delenv("XPFC_AUTOTUNE", raising=False)
    with Context(p) as ctx:
        ctx.set_field(psi0)
        free, total = _hip_mem_info()
        keep = total // 32    # less than the tuner's margin of total / 16
        filler = DeviceBuffer(ctx, max(free - keep, 1))
        try:
            assert _hip_mem_info()[0] < total // 16
            ctx.step(25)
            ctx.sync()
            tuned = ctx.variants()
        finally:
            filler.free()
        got = ctx.get_field()
        fe = [s.F_ref for s in ctx.samples() if s.has_fe]
    monkeypatch.setenv("XPFC_AUTOTUNE", "0")
    with Context(p) as ctx:
        ctx.set_variants(*tuned)
        ctx.set_field(psi0)
        ctx.step(25)
        ref = ctx.get_field()
        fe_ref = [s.F_ref for s in ctx.samples() if s.has_fe]
    with Context(p) as ctx:   # the default variants, untuned
        ctx.set_field(psi0)
        ctx.step(25)
        untuned = ctx.get_field()
    assert np.all(np.isfinite(got))
    assert np.array_equal(got, ref) and np.array_equal(got, untuned)
    assert fe == fe_ref and len(fe) >= 2
