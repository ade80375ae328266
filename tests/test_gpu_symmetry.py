"""GPU checks that the mathematics fixes, independent of any stored output:
lattice initial conditions against the oracle, axis-permutation symmetry on
cubic grids (the three axes run through three different kernels), translation
equivariance, a z-invariant 3-D run against the 2-D run, and the transforms
against a second library FFT (rocFFT, test-only)."""
import numpy as np
import pytest

from oracle import xpfc_oracle as O
from xpfc_cuda_amd import Context, XPFC_FP32_MIXED, XPFC_FP64

from cases import growth_case, make

pytestmark = pytest.mark.gpu

# absolute tolerances on fields of size ~0.05 +- 0.02 after a few steps
TOL = {XPFC_FP64: 1e-13, XPFC_FP32_MIXED: 1e-6}


def _flush(msg):
    print(msg, flush=True)


def _run(p, field, steps):
    with Context(p) as c:
        c.set_field(field)
        c.step(steps)
        return c.get_field()


@pytest.mark.parametrize("iconfig,iseed,ishape", [
    (0, 0, 0), (1, 0, 0), (2, 0, 0),          # triangular, square, liquid
    (0, 1, 0), (1, 1, 0),                     # circular seed of radius roa*a
    (0, 1, 1), (1, 1, 1),                     # Gaussian seed of width roa
])
@pytest.mark.parametrize("prec", [XPFC_FP64, XPFC_FP32_MIXED])
def test_lattice_ics_match_oracle(iconfig, iseed, ishape, prec):
    """initialize.cu:95-202: the library's xpfc_initialize against the oracle's
    restatement on a non-square box with an off-centre seed."""
    _flush("ic iconfig=%d iseed=%d ishape=%d prec=%d" % (iconfig, iseed, ishape, prec))
    op, p = make((64, 32), prec=prec)
    p.iconfig, p.iseed, p.ishape = iconfig, iseed, ishape
    p.A_10, p.A_11, p.A_tri = 0.1, 0.05, 0.2
    p.h, p.k, p.roa = 3.0, 2.5, 2.0
    want = O.lattice_ic(op, iconfig, A_10=0.1, A_11=0.05, A_tri=0.2, iseed=iseed,
                        ishape=ishape, h=3.0, kc=2.5, roa=2.0)
    with Context(p) as c:
        c.initialize()
        got = c.get_field()
        mean = c.avg_calc()
    # FP32: the fluctuation about the mean is transformed and held in float
    fluct = max(np.max(np.abs(want - p.psi0)), 1e-3)
    assert np.max(np.abs(got - want)) <= (1e-14 if prec == XPFC_FP64 else 1e-6 * fluct)
    assert mean == pytest.approx(p.psi0, abs=1e-14)


@pytest.mark.parametrize("n,prec,steps", [
    (32, XPFC_FP64, 10),
    (128, XPFC_FP32_MIXED, 5),   # 64 KiB planes: the padded plane stride
])
def test_cubic_axis_permutation(n, prec, steps):
    """On a cubic grid the step commutes with any permutation of the axes;
    axis 0 runs through the fused update pass, axis 1 through the strided pass
    and axis 2 through the real-space pass, so this ties the three together."""
    _flush("permutation %d^3 prec=%d" % (n, prec))
    _, p = growth_case((n, n, n), prec=prec, t_sample=0)
    rng = np.random.default_rng(11)
    x = 0.05 + 0.02 * rng.standard_normal((n, n, n))
    base = _run(p, x, steps)
    for perm in [(1, 0, 2), (0, 2, 1), (2, 1, 0), (1, 2, 0)]:
        got = _run(p, np.ascontiguousarray(x.transpose(perm)), steps)
        assert np.max(np.abs(got - base.transpose(perm))) <= TOL[prec], perm


@pytest.mark.parametrize("shape", [(64, 32), (16, 16, 32)])
def test_translation_equivariance(shape):
    """Shifting the initial field by one cell along any axis shifts the result."""
    _flush("translation %s" % (shape,))
    _, p = growth_case(shape, t_sample=0)
    rng = np.random.default_rng(12)
    x = 0.05 + 0.02 * rng.standard_normal(shape)
    base = _run(p, x, 12)
    for ax in range(len(shape)):
        got = _run(p, np.roll(x, 1, axis=ax), 12)
        assert np.max(np.abs(got - np.roll(base, 1, axis=ax))) <= TOL[XPFC_FP64], ax


@pytest.mark.parametrize("prec", [XPFC_FP64, XPFC_FP32_MIXED])
def test_z_invariant_3d_equals_2d(prec):
    """A field constant along z has only kz = 0 modes, so the 3-D step on
    (Mx, My, 8) reproduces the 2-D step on (Mx, My) in every z plane."""
    _flush("z-invariant prec=%d" % prec)
    _, p2 = growth_case((32, 64), prec=prec, t_sample=0)
    _, p3 = growth_case((32, 64, 8), prec=prec, t_sample=0)
    rng = np.random.default_rng(13)
    x = 0.05 + 0.02 * rng.standard_normal((32, 64))
    a = _run(p2, x, 20)
    b = _run(p3, np.repeat(x[:, :, None], 8, axis=2), 20)
    for z in range(8):
        assert np.max(np.abs(b[:, :, z] - a)) <= TOL[prec], z


@pytest.mark.parametrize("shape", [(256, 256), (64, 32, 128), (128, 128, 128)])
def test_fft_vs_second_library(shape):
    """The forward transform against rocFFT (called through ctypes) in FP64:
    two independent FFT implementations agree to FP64 roundoff."""
    import rocfft_ref
    _flush("fft vs rocFFT %s" % (shape,))
    _, p = make(shape, prec=XPFC_FP64)
    x = np.random.default_rng(14).standard_normal(shape)
    ref = rocfft_ref.rfftn(x)
    assert np.max(np.abs(ref - np.fft.rfftn(x))) <= 1e-12 * np.max(np.abs(ref)) * 4
    with Context(p) as c:
        got = c.fft_r2c(x)
    assert got.shape == ref.shape
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref)) * 4


@pytest.mark.parametrize("shape,mode", [((64, 32), (5, 3)), ((16, 32, 64), (3, 7, 9)),
                                        ((128, 128, 128), (17, 40, 21))])
@pytest.mark.parametrize("prec", [XPFC_FP64, XPFC_FP32_MIXED])
def test_fft_plane_wave_known_answer(shape, mode, prec):
    """R2C of cos(2 pi m.x / n) is N/2 in bin m of the half spectrum and zero
    elsewhere (the conjugate bin -m lies in the unstored half)."""
    _flush("plane wave %s %s prec=%d" % (shape, mode, prec))
    _, p = make(shape, prec=prec)
    grids = np.meshgrid(*[np.arange(n) / n for n in shape], indexing="ij")
    x = np.cos(2.0 * np.pi * sum(m * g for m, g in zip(mode, grids)))
    with Context(p) as c:
        got = c.fft_r2c(x)
    N = x.size
    expect = np.zeros_like(got)
    expect[mode] = N / 2.0
    tol = (1e-12 if prec == XPFC_FP64 else 2e-6) * N
    assert np.max(np.abs(got - expect)) <= tol
