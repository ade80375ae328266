"""Pins the FP64 oracle independently of its own formulas.

The reference ships no tests, fixtures or golden vectors, so each oracle
function is fixed here by mathematics: brute-force DFTs, closed-form evolution
of single Fourier modes, closed-form free energy / residual, conservation laws,
the 2-D <-> z-invariant 3-D reduction, and published Philox known answers.
"""
import math

import numpy as np
import pytest

from oracle import xpfc_oracle as O

TWO_PI = 2.0 * math.pi


def P(n=(32, 32), dx=0.125, **kw):
    kw.setdefault("peaks", [O.Peak(TWO_PI, 1.0, 1.0), O.Peak(TWO_PI * math.sqrt(2), 1.0, 1.0)])
    kw.setdefault("dcf_mode", "max")
    return O.Params(n=n, L=tuple(m * dx for m in n), **kw)


# ---------------------------------------------------------------- grid / FFT
def test_fft_freq_known_values():
    assert np.allclose(O.fft_freq(8, TWO_PI), [0, 1, 2, 3, 4, -3, -2, -1])  # Nyquist positive
    assert np.allclose(O.fft_freq(5, TWO_PI), [0, 1, 2, -2, -1])
    assert np.allclose(O.fft_freq(4, 2.0), np.array([0, 1, 2, -1]) * math.pi)


def test_linspace_endpoint_inclusive():
    v = O.linspace_ref(0.0, 4.0, 5)
    assert np.allclose(v, [0, 1, 2, 3, 4])


def test_numpy_fft_convention_brute_force():
    """cuFFT FORWARD is sum x e^{-2πi jk/N}, INVERSE unnormalised with +; the
    oracle uses numpy's fftn (= forward) and ifftn*N (= inverse)."""
    rng = np.random.default_rng(0)
    x = rng.standard_normal((4, 6)) + 1j * rng.standard_normal((4, 6))
    X = np.zeros_like(x)
    for k0 in range(4):
        for k1 in range(6):
            s = 0
            for j0 in range(4):
                for j1 in range(6):
                    s += x[j0, j1] * np.exp(-2j * np.pi * (j0 * k0 / 4 + j1 * k1 / 6))
            X[k0, k1] = s
    assert np.allclose(np.fft.fftn(x), X, atol=1e-12)
    back = np.zeros_like(x)
    for j0 in range(4):
        for j1 in range(6):
            back[j0, j1] = np.sum(X * np.exp(2j * np.pi * (j0 * np.arange(4)[:, None] / 4
                                                           + j1 * np.arange(6)[None, :] / 6)))
    assert np.allclose(np.fft.ifftn(X) * x.size, back, atol=1e-9)
    assert np.allclose(back, x * x.size, atol=1e-9)


def test_k_squared_plane_wave_is_laplacian():
    p = P(n=(16, 8), dx=0.25)
    k2 = O.k_squared(p)
    x = np.arange(16)[:, None] * 0.25
    y = np.arange(8)[None, :] * 0.25
    kx, ky = 3 * TWO_PI / p.L[0], 1 * TWO_PI / p.L[1]
    f = np.cos(kx * x + ky * y)
    lap = np.real(np.fft.ifftn(-k2 * np.fft.fftn(f)))
    assert np.allclose(lap, -(kx * kx + ky * ky) * f, atol=1e-10)


# ---------------------------------------------------------------- DCF
def test_dcf_peak_values_closed_form():
    sig = 0.1
    p = P(sigma=sig, peaks=[O.Peak(TWO_PI, 1.0, 4.0)])
    dw = math.exp(-(sig * TWO_PI) ** 2 / 8.0)
    v = O.c2_hat(np.array([TWO_PI, TWO_PI + 1.0, TWO_PI - 2.0, 0.0]), p)
    assert v[0] == pytest.approx(dw, rel=1e-15)
    assert v[1] == pytest.approx(dw * math.exp(-0.5), rel=1e-15)
    assert v[2] == pytest.approx(dw * math.exp(-2.0), rel=1e-15)
    assert v[3] == pytest.approx(dw * math.exp(-TWO_PI ** 2 / 2), rel=1e-12)


def test_dcf_sum_vs_max():
    k1, k2 = TWO_PI, TWO_PI * math.sqrt(2)
    for mode in ("sum", "max"):
        p = P(sigma=0.0, dcf_mode=mode, peaks=[O.Peak(k1, 1.0, 1.0), O.Peak(k2, 1.0, 1.0)])
        v = O.c2_hat(np.array([k1, k2]), p)
        cross = math.exp(-((k2 - k1) ** 2) / 2)
        if mode == "sum":
            assert v[0] == pytest.approx(1.0 + cross) and v[1] == pytest.approx(1.0 + cross)
        else:
            assert v[0] == pytest.approx(1.0) and v[1] == pytest.approx(1.0)
    # far apart peaks: sum == max to within exp(-large)
    p1 = P(sigma=0.0, dcf_mode="sum", peaks=[O.Peak(1.0, 0.1, 1.0), O.Peak(9.0, 0.1, 1.0)])
    p2 = P(sigma=0.0, dcf_mode="max", peaks=[O.Peak(1.0, 0.1, 1.0), O.Peak(9.0, 0.1, 1.0)])
    k = np.linspace(0, 10, 101)
    assert np.allclose(O.c2_hat(k, p1), O.c2_hat(k, p2), atol=1e-300, rtol=1e-12)


def test_reference_peak_constants():
    pk = O.reference_peaks()
    assert pk[0].k == TWO_PI and pk[0].rho_beta == 4.0
    assert pk[1].k == pytest.approx(TWO_PI * math.sqrt(2)) and pk[1].rho_beta == pytest.approx(4 / math.sqrt(2))


# ---------------------------------------------------------------- one mode, linear
def _mode(p, ix, iy, A, n0):
    x = np.arange(p.n[0])[:, None] * (p.L[0] / p.n[0])
    y = np.arange(p.n[1])[None, :] * (p.L[1] / p.n[1])
    kx, ky = ix * TWO_PI / p.L[0], iy * TWO_PI / p.L[1]
    return n0 + A * np.cos(kx * x + ky * y), kx * kx + ky * ky


@pytest.mark.parametrize("eps,M", [(0.0, 1.0), (0.1, 0.7)])
def test_linear_single_mode_closed_form(eps, M):
    """a = b = 0: A_n = A_0 / (1 + dt M k^2 (1 - eps - c2(k)))^n, mean fixed."""
    p = P(sigma=0.08, a=0.0, b=0.0, eps=eps, M=M, dt=0.5, t_sample=1000)
    psi, k2 = _mode(p, 8, 0, 0.01, 0.05)   # |k| = 2π: on the first peak
    c2 = math.exp(-(0.08 * TWO_PI) ** 2 / 2) * math.exp(-(math.sqrt(k2) - TWO_PI) ** 2 / 2)
    g = 1.0 / (1.0 + 0.5 * M * k2 * (1.0 - eps - c2))
    out, _ = O.run(psi, p, 25, sample=False)
    expect, _ = _mode(p, 8, 0, 0.01 * g ** 25, 0.05)
    assert np.allclose(out, expect, atol=1e-15)
    out2, _ = O.run(psi, p, 25, sample=False, literal=True)
    assert np.allclose(out2, expect, atol=1e-15)


def test_nonlinear_one_step_closed_form():
    """psi = A cos(kx): P2 = A²/2 (1 + cos 2kx), P3 = A³ (3 cos kx + cos 3kx)/4,
    so one step gives modes k, 2k, 3k with explicit amplitudes."""
    a, b, dt, M = 0.7, 0.4, 0.3, 1.0
    p = P(n=(64, 8), sigma=0.1, a=a, b=b, dt=dt, M=M, t_sample=1000)
    A = 0.2
    psi, k2 = _mode(p, 2, 0, A, 0.0)
    k = math.sqrt(k2)

    def den(kk):
        c2 = O.c2_hat(np.array([kk]), p)[0]
        return 1.0 + dt * M * kk * kk * (1.0 - c2)

    amp1 = (A - b * k2 * M * dt * 3 * A ** 3 / 4) / den(k)
    amp2 = (a * (2 * k) ** 2 * M * dt * A ** 2 / 2) / den(2 * k)
    amp3 = (-b * (3 * k) ** 2 * M * dt * A ** 3 / 4) / den(3 * k)
    x = np.arange(64)[:, None] * (p.L[0] / 64)
    expect = amp1 * np.cos(k * x) + amp2 * np.cos(2 * k * x) + amp3 * np.cos(3 * k * x)
    expect = np.broadcast_to(expect, p.n)
    out, _ = O.run(psi, p, 1, sample=False)
    assert np.allclose(out, expect, atol=1e-15)


# ---------------------------------------------------------------- FE, mean, residual
def test_fe_constant_field_closed_form():
    p = P(n=(16, 32), sigma=0.1, a=1.0, b=1 / 3, dcf_mode="sum", peaks=O.reference_peaks())
    n0 = 0.07
    psi = np.full(p.n, n0)
    c20 = O.c2_hat(np.array([0.0]), p)[0]
    f = n0 ** 2 / 2 - n0 ** 3 / 3 + (1 / 3) * n0 ** 4 / 4 - n0 ** 2 * c20 / 2
    ff = p.L[0] * p.L[1] / ((p.n[0] - 1) * (p.n[1] - 1))
    F, fm = O.free_energy(psi, p, O.k_tables(p)[1])
    assert fm == pytest.approx(f, rel=1e-13)
    assert F == pytest.approx(ff * f * psi.size, rel=1e-13)


def test_fe_single_mode_closed_form():
    a, b = 1.0, 1.0 / 3.0
    p = P(n=(32, 16), sigma=0.1, a=a, b=b)
    n0, A = 0.05, 0.03
    psi, k2 = _mode(p, 4, 2, A, n0)
    N = psi.size
    c2k = O.c2_hat(np.array([math.sqrt(k2)]), p)[0]
    c20 = O.c2_hat(np.array([0.0]), p)[0]
    s2 = n0 ** 2 + A ** 2 / 2
    s3 = n0 ** 3 + 1.5 * n0 * A ** 2
    s4 = n0 ** 4 + 3 * n0 ** 2 * A ** 2 + 3 * A ** 4 / 8
    conv = c20 * n0 ** 2 + c2k * A ** 2 / 2
    f = 0.5 * s2 - a / 3 * s3 + b / 4 * s4 - 0.5 * conv
    F, fm = O.free_energy(psi, p, O.k_tables(p)[1])
    assert fm == pytest.approx(f, rel=1e-12)
    assert F == pytest.approx(f * N * p.L[0] * p.L[1] / (31 * 15), rel=1e-12)


@pytest.mark.parametrize("M,int_Mo", [(1.0, True), (1.5, True), (1.5, False)])
def test_residual_linear_closed_form(M, int_Mo):
    """a = b = 0, eps = 0: R = (g-1)/dt A + M_r k²(1-c2) A with g the growth factor."""
    dt = 0.5
    p = P(sigma=0.08, a=0.0, b=0.0, M=M, dt=dt, int_Mo=int_Mo, t_sample=1)
    A = 0.01
    psi, k2 = _mode(p, 8, 0, A, 0.05)
    c2 = O.c2_hat(np.array([math.sqrt(k2)]), p)[0]
    g = 1.0 / (1.0 + dt * M * k2 * (1.0 - c2))
    Mr = float(int(M)) if int_Mo else M
    r_amp = (g - 1.0) / dt * A + Mr * k2 * (1.0 - c2) * A
    _, s = O.run(psi, p, 1, sample=True)
    rmse = [x.value for x in s if x.kind == "SSE" and x.step == 0][0]
    assert rmse == pytest.approx(abs(r_amp) / math.sqrt(2), rel=1e-9)


def test_mass_conservation_and_schedule():
    p = P(sigma=0.07, t_sample=10)
    psi = O.noise_ic(p, 0.01, 1)
    out, s = O.run(psi, p, 40, sample=True)
    assert abs(np.mean(out) - 0.05) < 1e-16
    kinds = [(x.step, x.kind) for x in s]
    # FE at start, SSE at n % 10 == 0, FE at (n+1) % 10 == 0, FE at end (main.cu:286,326,412,450)
    assert kinds[:3] == [(0, "FE"), (0, "SSE"), (10, "FE")]
    assert kinds[-2:] == [(40, "FE"), (40, "FE")]
    assert sum(1 for k in kinds if k[1] == "SSE") == 4


def test_literal_complex_path_equals_real_path_nonlinear():
    p = P(n=(16, 16), sigma=0.07, t_sample=5)
    psi = O.noise_ic(p, 0.02, 2)
    a, sa = O.run(psi, p, 15, sample=True, literal=False)
    b, sb = O.run(psi, p, 15, sample=True, literal=True)
    assert np.max(np.abs(a - b)) < 1e-15
    for x, y in zip(sa, sb):
        assert x.step == y.step and x.kind == y.kind
        assert x.value == pytest.approx(y.value, rel=1e-11, abs=1e-300)


def test_3d_z_invariant_reduces_to_2d():
    p2 = P(n=(16, 32), sigma=0.07, t_sample=1000)
    p3 = P(n=(16, 32, 4), sigma=0.07, t_sample=1000)
    psi2 = O.noise_ic(p2, 0.02, 3)
    psi3 = np.repeat(psi2[:, :, None], 4, axis=2)
    o2, _ = O.run(psi2, p2, 10, sample=False)
    o3, _ = O.run(psi3, p3, 10, sample=False)
    assert np.max(np.abs(o3 - o2[:, :, None])) < 1e-15


def test_translation_equivariance():
    p = P(n=(16, 16), sigma=0.07)
    psi = O.noise_ic(p, 0.02, 4)
    a, _ = O.run(psi, p, 10, sample=False)
    b, _ = O.run(np.roll(psi, (3, 5), axis=(0, 1)), p, 10, sample=False)
    assert np.max(np.abs(np.roll(a, (3, 5), axis=(0, 1)) - b)) < 1e-15


def test_powf_fp32_switch_rounds_to_float():
    x = np.array([0.1, 0.05, -0.3333333333333])
    got = O.power_function(x, 3, powf_fp32=True)
    xf = x.astype(np.float32)
    assert np.array_equal(got, ((xf * xf).astype(np.float32) * xf).astype(np.float32).astype(np.float64))
    assert not np.array_equal(got, x ** 3)


# ---------------------------------------------------------------- RNG and ICs
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_philox_kat_oracle(ctr, key, expect):
    out = O.philox4x32_10(np.array([ctr], dtype=np.uint64), key)[0]
    assert tuple(int(v) for v in out) == expect


def test_normals_statistics():
    z = O.philox_normals(1 << 20, 99)
    assert abs(z.mean()) < 5e-3 and abs(z.var() - 1.0) < 5e-3
    assert abs(np.mean(z ** 3)) < 1e-2 and abs(np.mean(z ** 4) - 3.0) < 3e-2


def test_noise_ic_mean_fixed():
    p = P(n=(32, 16), psi0=0.04)
    psi = O.noise_ic(p, 0.01, 7)
    assert abs(psi.mean() - 0.04) < 1e-17
    assert 0.009 < psi.std() < 0.011


def test_square_ic_values():
    p = P(n=(8, 8), dx=0.25, psi0=0.05)
    psi = O.lattice_ic(p, 1, A_10=0.1, A_11=0.05)
    x = O.linspace_ref(0.0, 2.0, 8)
    raw = (0.1 * ((3 * math.sqrt(2) - 2) * np.cos(TWO_PI * x)[None, :] - 2 * np.cos(TWO_PI * x)[:, None])
           - 0.05 * np.cos(TWO_PI * x)[:, None] * np.cos(TWO_PI * x)[None, :]) + 0.05
    shift = 0.05 - raw.mean()
    assert np.allclose(psi, raw + shift, atol=1e-16)
    assert abs(psi.mean() - 0.05) < 1e-16


# ---------------------------------------------------------------- lattice ICs
# lattice_ic returns the field after the mean fix, which adds one constant to
# every cell: the checks below use differences between cells, or cells whose
# raw value is known exactly, so they hold whatever that constant is.
S3 = math.sqrt(3.0)


def _box(nx, ny, Lx, Ly, psi0=0.05):
    return O.Params(n=(nx, ny), L=(Lx, Ly), psi0=psi0)


def test_triangular_lattice_special_points_and_mirrors():
    """Triangular one-mode field A(cos(qx)/2 + cos(qx/2)cos(sqrt3 qy/2)), q = 2pi/a_tri
    = sqrt3 pi (initialize.cu:116-118).  On x_i = i/(2 sqrt3), y_j = j/6 the
    phases are qx = i pi/2 and sqrt3 qy/2 = j pi/4, so (by hand): all three
    waves in phase at the origin, 3A/2; -A/2 at (0, 4), (4, 0), (2, 0), (2, 2);
    A/sqrt2 at (1, 0); and the field is even about the box centre in x and y."""
    A = 0.3
    p = _box(9, 9, 4.0 / S3, 4.0 / 3.0)
    psi = O.lattice_ic(p, 0, A_tri=A)
    d = psi - psi[0, 0] + 1.5 * A          # raw deviation, fixed at the origin's 3A/2
    for (i, j) in [(0, 4), (4, 0), (2, 0), (2, 2), (6, 2)]:
        assert d[i, j] == pytest.approx(-0.5 * A, abs=1e-14), (i, j)
    assert d[8, 8] == pytest.approx(1.5 * A, abs=1e-14)
    assert d[1, 0] == pytest.approx(A / math.sqrt(2.0), abs=1e-14)   # cos(pi/2)/2 + cos(pi/4)
    assert d[0, 1] == pytest.approx(A * (0.5 + 1.0 / math.sqrt(2.0)), abs=1e-14)
    assert np.allclose(psi, psi[::-1, :], atol=1e-14)   # mirror in x about Lx/2
    assert np.allclose(psi, psi[:, ::-1], atol=1e-14)   # mirror in y about Ly/2
    assert abs(psi.mean() - 0.05) < 1e-16


def test_square_lattice_special_points():
    """Square field A10((3sqrt2-2)cos(2pi y) - 2cos(2pi x)) - A11 cos(2pi x)cos(2pi y)
    (initialize.cu:96-99) on x_i = i/4, y_j = j/4 (cosines 1, 0, -1, 0, 1):
    by hand, psi(2,0)-psi(0,0) = 4A10 + 2A11, psi(0,2)-psi(0,0) = -2(3sqrt2-2)A10
    + 2A11, psi(1,1)-psi(0,0) = -(3sqrt2-4)A10 + A11.  Swapping x and y, or the
    sign of either amplitude, changes one of them."""
    A10, A11 = 0.07, 0.03
    psi = O.lattice_ic(_box(5, 5, 1.0, 1.0), 1, A_10=A10, A_11=A11)
    r = 3.0 * math.sqrt(2.0)
    assert psi[2, 0] - psi[0, 0] == pytest.approx(4 * A10 + 2 * A11, abs=1e-15)
    assert psi[0, 2] - psi[0, 0] == pytest.approx(-2 * (r - 2) * A10 + 2 * A11, abs=1e-15)
    assert psi[1, 1] - psi[0, 0] == pytest.approx(-(r - 4) * A10 + A11, abs=1e-15)
    assert psi[4, 4] == pytest.approx(psi[0, 0], abs=1e-15)
    assert abs(psi.mean() - 0.05) < 1e-16


@pytest.mark.parametrize("iconfig,a_lat", [(0, 2.0 / S3), (1, 1.0)])
def test_circle_seed(iconfig, a_lat):
    """initialize.cu:158-172: cells farther than r = roa * a (a = a_tri or a_sq)
    from (h, k) are psi0; nearer cells keep the lattice value.  After the mean
    fix: every outside cell holds one value, and inside cells differ from the
    unseeded field by one constant.  Cells with roa < d <= roa*a_tri must be
    inside for the triangular lattice (the radius uses the lattice constant)."""
    p = _box(33, 33, 4.0, 4.0)
    roa, h, kc = 1.0, 2.0, 2.0
    kw = dict(A_10=0.07, A_11=0.03, A_tri=0.2)
    seeded = O.lattice_ic(p, iconfig, iseed=1, ishape=0, h=h, kc=kc, roa=roa, **kw)
    plain = O.lattice_ic(p, iconfig, **kw)
    x = np.arange(33) * (4.0 / 32)
    d = np.sqrt((x[:, None] - h) ** 2 + (x[None, :] - kc) ** 2)
    out, ins = d > roa * a_lat, d <= roa * a_lat
    assert np.ptp(seeded[out]) < 1e-15
    diff = seeded[ins] - plain[ins]
    assert np.ptp(diff) < 1e-14
    if iconfig == 1:
        # the centre cell (x = y = 2, all cosines 1) has the raw value
        # A10(3sqrt2 - 4) - A11 + psi0, which fixes the seeded field's shift:
        # the outside cells then hold psi0 before it
        raw_c = 0.07 * (3.0 * math.sqrt(2.0) - 4.0) - 0.03 + 0.05
        shift = seeded[16, 16] - raw_c
        assert seeded[out][0] - shift == pytest.approx(0.05, abs=1e-15)
    ring = (d > roa) & (d <= roa * a_lat)
    if iconfig == 0:
        assert ring.sum() > 0
        assert np.ptp(seeded[ring] - plain[ring]) < 1e-14   # lattice values kept there


def test_gaussian_seed():
    """initialize.cu:174-188: psi <- (psi - psi0) G + psi0 (1 - G), G =
    exp(-|r - (h,k)|^2 / roa) (roa is the variance, not its square root).  On
    the triangular grid of the special-points test, 8 periods wide, with the
    centre on a lattice maximum: far cells (G = 0 exactly) hold psi0 before the
    shift; the centre holds the lattice deviation 3A/2 (not 3A/2 + psi0); the
    next cell along x holds (A/sqrt2) G + psi0 (1 - G) with G = exp(-1/(12 roa))."""
    A, psi0, roa = 0.2, 0.05, 0.05
    p = _box(65, 65, 32.0 / S3, 32.0 / 3.0, psi0=psi0)
    x = np.arange(65) * (p.L[0] / 64)
    y = np.arange(65) * (p.L[1] / 64)
    h, kc = x[32], y[32]
    psi = O.lattice_ic(p, 0, A_tri=A, iseed=1, ishape=1, h=h, kc=kc, roa=roa)
    far = psi[0, 0]
    shift = far - psi0
    assert psi[64, 64] == far and psi[0, 64] == far
    assert psi[32, 32] - shift == pytest.approx(1.5 * A, abs=1e-14)
    G = math.exp(-1.0 / (12.0 * roa))
    assert psi[33, 32] - shift == pytest.approx(A / math.sqrt(2.0) * G + psi0 * (1 - G), abs=1e-14)
    assert psi[31, 32] == pytest.approx(psi[33, 32], abs=1e-14)


# ---------------------------------------------------------------- powf_fp32 bound
POWF_ULP = 4   # CUDA's documented maximum error of powf (full range), in ulp


@pytest.mark.parametrize("pw", [2, 3, 4])
def test_powf_fp32_within_cuda_documented_bound(pw):
    """power_functions.cu:47 computes powf((float)psi, p).  The oracle's switch
    forms float products of float(psi); each result must be a float and lie
    within POWF_ULP ulp of the exact power of float(psi) (evaluated with exact
    rational arithmetic), the bound CUDA documents for powf itself; so the
    switch and CUDA's powf differ by at most twice that."""
    from fractions import Fraction
    rng = np.random.default_rng(pw)
    x = np.concatenate([rng.uniform(-1.5, 1.5, 2000), rng.uniform(-0.1, 0.1, 2000),
                        [0.0, 1.0, -1.0, 0.05, 1e-20]])
    got = O.power_function(x, pw, powf_fp32=True)
    assert np.array_equal(got, got.astype(np.float32).astype(np.float64))  # FP32 values
    xf = x.astype(np.float32)
    worst = 0.0
    for g, v in zip(got, xf):
        exact = Fraction(float(v)) ** pw
        ex32 = np.float32(float(exact))
        ulp = float(np.spacing(np.abs(ex32))) if ex32 != 0 else float(np.finfo(np.float32).tiny)
        err = abs(Fraction(float(g)) - exact) / Fraction(ulp)
        worst = max(worst, float(err))
    assert worst <= POWF_ULP, worst
    # the slip of not rounding the input to float first is visible
    assert not np.array_equal(got, (x ** pw).astype(np.float32).astype(np.float64))


# ---------------------------------------------------------------- divergence
def _square(n, sigma, psi0):
    return O.Params(n=n, L=tuple(m * 0.125 for m in n), sigma=sigma, psi0=psi0,
                    peaks=[O.Peak(TWO_PI, 1.0, 1.0), O.Peak(TWO_PI * math.sqrt(2), 1.0, 1.0)],
                    dcf_mode="max", t_sample=10 ** 6)


def _blowup(p, amp, seed, nmax, record=()):
    st = O.Stepper(O.noise_ic(p, amp, seed), p)
    amp_at = {}
    with np.errstate(over="ignore", invalid="ignore"):
        for s in range(1, nmax + 1):
            st.step()
            a = float(np.max(np.abs(st.psi - p.psi0)))
            if s in record:
                amp_at[s] = a
            if not np.isfinite(a):
                return s, amp_at
    return None, amp_at


def test_oracle_diverges_on_the_strongly_crystallising_job():
    """The service test's job 4 (64^2, sigma 0.062, n0 0.0564, noise 0.01 seed 4)
    diverges under the reference's explicit nonlinear term in FP64 too: the
    amplitude is 0.03 at step 10, 0.13 at 15, 1.3 at 19 and the field is
    non-finite before step 40 (well inside the job's 137 steps)."""
    s, a = _blowup(_square((64, 64), 0.062, 0.0564), 0.01, 4, 137, record=(10, 15, 19))
    assert s is not None and s < 40
    assert 0.02 < a[10] < 0.05 and 0.1 < a[15] < 0.2 and 1.0 < a[19] < 2.0


def test_oracle_diverges_on_the_slow_growth_case():
    """The sigma = 0.07 growth case (32^2, n0 0.05, seed 1) used by the GPU
    parity tests for up to 80 steps: amplitude 0.002 at step 20 and 0.04 at
    step 100, non-finite shortly after (between 100 and 160)."""
    s, a = _blowup(_square((32, 32), 0.07, 0.05), 0.01, 1, 160, record=(20, 80, 100))
    assert s is not None and 100 < s <= 160
    assert a[20] < 0.003 and a[80] < 0.03 and 0.02 < a[100] < 0.1


def test_vtk_layout_x_fastest():
    psi = np.arange(6, dtype=float).reshape(3, 2)  # nx = 3, ny = 2
    t = O.vtk_text(psi, 3, 4).splitlines()
    assert t[4] == "DIMENSIONS 3 2 1" and t[6] == "SPACING 1 2 1" and t[7] == "POINT_DATA 6"
    assert [float(v) for v in t[10:]] == [0, 2, 4, 1, 3, 5]
