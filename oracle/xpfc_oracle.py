"""FP64 host oracle for the XPFC time-stepping path.

TEST INFRASTRUCTURE ONLY.  This module is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product (``xpfc_cuda_amd`` and its HIP library) never calls it,
and never falls back to it.

It restates, in numpy (FP64, pocketfft via ``numpy.fft``/``scipy.fft``), the
equations of LC2500/XPFC_CUDA as they are coded.  Every function names the
reference ``file:line`` it follows.  The reference is 2-D only; the 3-D form
used here adds the third axis to ``k^2`` and to the free-energy normalisation
and changes nothing else.

Quirk switches (all default to the reference's behaviour where it matters):

* ``powf_fp32``   -- ψ^p evaluated as ``float(powf(float ψ, p))``
                     (power_functions.cu:47; reused for the residual square,
                     sse_calculation.cu:38).
* ``int_Mo``      -- the residual truncates the mobility to int
                     (sse_calculation.cu:120, globals.cuh:111).
* ``dcf_mode``    -- "sum" (initialize.cu:215-217, active) or "max"
                     (initialize.cu:209-212, the envelope form).

Parity pinning (what fixes each function independently of this file) lives in
``tests/test_oracle.py``: brute-force DFTs, closed-form single-mode evolution,
closed-form free energy and residual, conservation laws and the Philox
known-answer vectors.  Functions with no such check: none.

Parity unpinned (these live in closed libraries outside the reference and no
reference run is possible here): the exact cuFFT rounding sequence, the exact
CUDA ``powf`` FP32 outputs (``powf_fp32`` is held only to CUDA's documented
error bound, not to its bits), and the reference's printed F / RMSE values on
any particular input.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Sequence, Tuple

import numpy as np

try:  # pocketfft with threads when available
    import scipy.fft as _sfft

    def _rfftn(x, workers):
        return _sfft.rfftn(x, workers=workers)

    def _irfftn(x, s, workers):
        return _sfft.irfftn(x, s=s, workers=workers)

    def _fftn(x, workers):
        return _sfft.fftn(x, workers=workers)

    def _ifftn(x, workers):
        return _sfft.ifftn(x, workers=workers)
except Exception:  # pragma: no cover - numpy fallback
    def _rfftn(x, workers):
        return np.fft.rfftn(x)

    def _irfftn(x, s, workers):
        return np.fft.irfftn(x, s=s)

    def _fftn(x, workers):
        return np.fft.fftn(x)

    def _ifftn(x, workers):
        return np.fft.ifftn(x)


PI = math.pi


# ----------------------------------------------------------------------------
# parameters
# ----------------------------------------------------------------------------
@dataclass
class Peak:
    """One DCF peak: position k_i, width alpha_i and Debye-Waller rho_i*beta_i.

    initialize.cu:10-17 holds k_10=2π, k_11=2π√2, rho_10=1, rho_11=1/√2,
    B_10=B_11=4, al_10=al_11=1; the factor used is
    exp(-(σ k_i)^2 / (2 rho_i B_i)) (initialize.cu:216).
    """
    k: float
    alpha: float = 1.0
    rho_beta: float = 1.0


def reference_peaks() -> List[Peak]:
    """The two peaks hard-coded at initialize.cu:10-17."""
    return [Peak(2.0 * PI, 1.0, 1.0 * 4.0),
            Peak(2.0 * PI * math.sqrt(2.0), 1.0, (1.0 / math.sqrt(2.0)) * 4.0)]


@dataclass
class Params:
    n: Tuple[int, ...]            # (Mx, My) or (Mx, My, Mz); last axis contiguous
    L: Tuple[float, ...]          # box lengths (Lx, Ly[, Lz])
    dt: float = 0.5
    M: float = 1.0                # mobility Mo
    eps: float = 0.0
    a: float = 1.0                # a = 1/a_in   (read_input.cu:28)
    b: float = 1.0 / 3.0          # b = 1/b_in   (read_input.cu:28)
    sigma: float = 0.1
    peaks: List[Peak] = field(default_factory=reference_peaks)
    dcf_mode: str = "sum"
    psi0: float = 0.05
    t_sample: int = 100
    powf_fp32: bool = False
    int_Mo: bool = True

    @property
    def dim(self) -> int:
        return len(self.n)

    @property
    def ncells(self) -> int:
        return int(np.prod(self.n))


def params_from_eta_chi(eta: float, chi: float) -> Tuple[float, float]:
    """f = n^2/2 - eta n^3/6 + chi n^4/12 equals FE_calc.cu:64's
    P2/2 - (a/3) P3 + (b/4) P4 with a = eta/2, b = chi/3."""
    return eta / 2.0, chi / 3.0


# ----------------------------------------------------------------------------
# grid, wavenumbers, DCF
# ----------------------------------------------------------------------------
def fft_freq(N: int, L: float) -> np.ndarray:
    """FFT_freq.cu:4-26: k_i = 2π/L * i for i <= N/2 (even N: Nyquist positive),
    else 2π/L * (i - N).  The reference's one-past-the-end write is not kept."""
    k = np.empty(N, dtype=np.float64)
    if N % 2:
        for i in range(N):
            k[i] = (2.0 * PI / L) * (i if i < (N + 1) // 2 else i - N)
    else:
        for i in range(N):
            k[i] = (2.0 * PI / L) * (i if i < (N + 2) // 2 else i - N)
    return k


def linspace_ref(x1: float, x2: float, N: int) -> np.ndarray:
    """linspace.cu:5-18: vec[0]=x1, vec[i+1]=vec[i]+x2/(N-1) (endpoint
    inclusive when x1=0); the one-past-the-end write is not kept."""
    du = x2 / (N - 1)
    v = np.empty(N, dtype=np.float64)
    v[0] = x1
    for i in range(N - 1):
        v[i + 1] = v[i] + du
    return v


def k_squared(p: Params) -> np.ndarray:
    """initialize.cu:77-85: k2 = kx[i]^2 + ky[j]^2 (+ kz[l]^2 in 3-D)."""
    ks = [fft_freq(n, L) for n, L in zip(p.n, p.L)]
    grids = np.meshgrid(*ks, indexing="ij")
    k2 = np.zeros(p.n, dtype=np.float64)
    for g in grids:
        k2 = k2 + g * g
    return k2


def c2_hat(kmag: np.ndarray, p: Params) -> np.ndarray:
    """initialize.cu:209-217: sum (active) or max (envelope) over peaks of
    exp(-(σ k_i)^2 / (2 rho_i B_i)) * exp(-(k - k_i)^2 / (2 alpha_i^2))."""
    terms = [math.exp(-(p.sigma * pk.k) ** 2 / (2.0 * pk.rho_beta))
             * np.exp(-((kmag - pk.k) ** 2) / (2.0 * pk.alpha ** 2))
             for pk in p.peaks]
    if p.dcf_mode == "sum":
        out = np.zeros_like(kmag)
        for t in terms:
            out = out + t
        return out
    if p.dcf_mode == "max":
        out = terms[0]
        for t in terms[1:]:
            out = np.maximum(out, t)
        return out
    raise ValueError("dcf_mode must be 'sum' or 'max'")


def k_tables(p: Params) -> Tuple[np.ndarray, np.ndarray]:
    """(k2, c2_k) over the full grid, as initialize.cu:77-85 and :222-229."""
    k2 = k_squared(p)
    return k2, c2_hat(np.sqrt(k2), p)


# ----------------------------------------------------------------------------
# operators (one function per reference kernel)
# ----------------------------------------------------------------------------
def power_function(psi: np.ndarray, pw: int, powf_fp32: bool = False) -> np.ndarray:
    """power_functions.cu:34-53: out = (powf(psi.x, p), 0).  With powf_fp32 the
    input is rounded to float and the product formed in float (CUDA's powf is
    documented to within 2 ulp for integer exponents; no bit claim)."""
    x = np.real(psi)
    if powf_fp32:
        xf = x.astype(np.float32)
        r = np.ones_like(xf)
        for _ in range(pw):
            r = (r * xf).astype(np.float32)
        return r.astype(np.float64)
    return x ** pw


def euler_step(psi_k, p2_k, p3_k, c2_k, k2, p: Params):
    """euler_step.cu:50-59:
    psi_k' = (psi_k + a k2 M dt P2_k - b k2 dt M P3_k)
             / (1 + k2 dt M (1 - eps) - k2 M dt c2_k)."""
    num = psi_k + p.a * k2 * p.M * p.dt * p2_k - p.b * k2 * p.dt * p.M * p3_k
    den = 1.0 + k2 * p.dt * p.M * (1.0 - p.eps) - k2 * p.M * p.dt * c2_k
    return num / den


def scaling_step(f: np.ndarray, M2: int) -> np.ndarray:
    """scaling_step.cu:48-49: f /= M2."""
    return f / float(M2)


def copy_matrix(f: np.ndarray) -> np.ndarray:
    """copy_matrix.cu:44-45: out = (Re f, 0)."""
    return np.real(f).astype(np.complex128)


def convolve_fields(psi_k, c2_k):
    """convolve_fields.cu:41-42: convl = c2_k * psi_k."""
    return c2_k * psi_k


def fe_density(psi, p2, p3, p4, conv_r, p: Params) -> np.ndarray:
    """FE_calc.cu:64: f = 0.5 P2 - (a/3) P3 + (b/4) P4 - 0.5 psi conv."""
    return (0.5 * p2 - (p.a / 3.0) * p3 + (p.b / 4.0) * p4
            - 0.5 * np.real(psi) * np.real(conv_r))


def ff_factor(p: Params) -> float:
    """initialize.cu:63: Lx*Ly / ((Mx-1)(My-1)); 3-D adds Lz/(Mz-1)."""
    num = 1.0
    den = 1.0
    for n, L in zip(p.n, p.L):
        num *= L
        den *= (n - 1)
    return num / den


def residual_calc(psi_kp1, psi_k, p2_k, p3_k, c2_k, k2, p: Params):
    """sse_calculation.cu:133-140 with Mo passed as int (:120) when int_Mo:
    res = (psi_kp1 - psi_k)/dt + Mo k2 (psi_k - a P2 + b P3 - c2 psi_k)."""
    Mo = float(int(p.M)) if p.int_Mo else p.M
    d2mu = -Mo * k2 * (psi_k - p.a * p2_k + p.b * p3_k - c2_k * psi_k)
    return (psi_kp1 - psi_k) / p.dt - d2mu


def mean(psi: np.ndarray) -> float:
    """Avg_calc.cu:3-18: host mean of psi."""
    return float(np.sum(np.real(psi)) / np.real(psi).size)


def configure_avg_density(psi: np.ndarray, psi0: float) -> np.ndarray:
    """Avg_calc.cu:21-43: shift psi so that its mean equals psi_0."""
    m = mean(psi)
    diff = psi0 - m if m != psi0 else 0.0
    return np.real(psi) + diff


# ----------------------------------------------------------------------------
# diagnostics composed as main.cu composes them
# ----------------------------------------------------------------------------
def free_energy(psi_real: np.ndarray, p: Params, c2_k: np.ndarray,
                workers: int = 1) -> Tuple[float, float]:
    """main.cu:414-426 + FE_calc.cu:3-68 + sum_reduction.cu:56-66.

    Returns (F_ref, f_mean): F_ref = ff_factor * Σ f is the value the reference
    prints; f_mean = Σ f / M2."""
    psi_k = _fftn(psi_real.astype(np.complex128), workers)
    conv = scaling_step(_ifftn(convolve_fields(psi_k, c2_k), workers) * psi_real.size,
                        psi_real.size)
    p2 = power_function(psi_real, 2, p.powf_fp32)
    p3 = power_function(psi_real, 3, p.powf_fp32)
    p4 = power_function(psi_real, 4, p.powf_fp32)
    f = fe_density(psi_real, p2, p3, p4, conv, p)
    s = float(np.sum(f))
    return s * ff_factor(p), s / psi_real.size


def rmse_from_residual(res_k: np.ndarray, p: Params, workers: int = 1) -> float:
    """sse_calculation.cu:33-38 + sum_reduction.cu:69-80:
    sqrt( Σ_x square(Re IFFT(res)/M2) / M2 ), square via powf when powf_fp32."""
    M2 = res_k.size
    r = np.real(_ifftn(res_k, workers) * M2) / M2
    sq = power_function(r, 2, p.powf_fp32)
    return math.sqrt(float(np.sum(sq)) / M2)


# ----------------------------------------------------------------------------
# the time loop
# ----------------------------------------------------------------------------
@dataclass
class Sample:
    step: int
    kind: str        # "FE" or "SSE"
    value: float     # F_ref for FE, RMSE for SSE
    f_mean: float = float("nan")
    mean: float = float("nan")


class Stepper:
    """The real-transform form of the loop body of main.cu:309-408 with the
    k-space tables built once: psi_k' = euler_step(psi_k, FFT P2, FFT P3),
    psi = IFFT(psi_k')/M2, psi_k = FFT(psi)."""

    def __init__(self, psi_real: np.ndarray, p: Params, workers: int = 1):
        self.p = p
        self.workers = workers
        k2, c2 = k_tables(p)
        h = p.n[-1] // 2 + 1
        self.c2_full = c2
        self.k2 = np.ascontiguousarray(k2[..., :h])
        self.c2 = np.ascontiguousarray(c2[..., :h])
        self.psi = np.array(psi_real, dtype=np.float64)
        self.psi_k = _rfftn(self.psi, workers)

    def step(self, want_sse: bool = False):
        p, w = self.p, self.workers
        p2_k = _rfftn(power_function(self.psi, 2, p.powf_fp32), w)
        p3_k = _rfftn(power_function(self.psi, 3, p.powf_fp32), w)
        psi_kp1 = euler_step(self.psi_k, p2_k, p3_k, self.c2, self.k2, p)
        rmse = None
        if want_sse:
            res = residual_calc(psi_kp1, self.psi_k, p2_k, p3_k, self.c2, self.k2, p)
            rmse = rmse_from_residual(_full_from_half(res, p.n, w), p, w)
        self.psi = _irfftn(psi_kp1, tuple(p.n), w)
        self.psi_k = _rfftn(self.psi, w)
        return rmse


def run(psi0_real: np.ndarray, p: Params, nsteps: int, sample: bool = True,
        workers: int = 1, literal: bool = False):
    """The time loop of main.cu:286-455 on one device.

    Per step n (main.cu:306-435):
      P2, P3 of psi_n -> FFT -> euler_step -> [SSE if n % t_sample == 0]
      -> IFFT, /M2, Re -> psi_{n+1} -> FFT -> [FE if (n+1) % t_sample == 0].
    FE is also taken at start (main.cu:286) and at the end (main.cu:450).

    ``literal=True`` uses full complex FFTs exactly as the reference; otherwise
    real-input transforms are used, which compute the same numbers in exact
    arithmetic because the reference keeps only Re(IFFT) (copy_matrix.cu:44-45)
    and its spectra are therefore Hermitian (pinned by a test).
    Returns (psi_final_real, samples)."""
    assert psi0_real.shape == tuple(p.n)
    M2 = p.ncells
    samples: List[Sample] = []
    if not literal:
        st = Stepper(psi0_real, p, workers)
        if sample:
            F, fm = free_energy(st.psi, p, st.c2_full, workers)
            samples.append(Sample(0, "FE", F, fm, mean(st.psi)))
        for n in range(nsteps):
            rmse = st.step(want_sse=sample and n % p.t_sample == 0)
            if rmse is not None:
                samples.append(Sample(n, "SSE", rmse))
            if sample and (n + 1) % p.t_sample == 0:
                F, fm = free_energy(st.psi, p, st.c2_full, workers)
                samples.append(Sample(n + 1, "FE", F, fm, mean(st.psi)))
        if sample:
            F, fm = free_energy(st.psi, p, st.c2_full, workers)
            samples.append(Sample(nsteps, "FE", F, fm, mean(st.psi)))
        return st.psi, samples
    k2, c2 = k_tables(p)
    psi = np.array(psi0_real, dtype=np.float64)
    if sample:
        F, fm = free_energy(psi, p, c2, workers)
        samples.append(Sample(0, "FE", F, fm, mean(psi)))
    # literal restatement: full complex transforms as the reference
    psi_k = _fftn(psi.astype(np.complex128), workers)
    for n in range(nsteps):
        p2_k = _fftn(power_function(psi, 2, p.powf_fp32).astype(np.complex128), workers)
        p3_k = _fftn(power_function(psi, 3, p.powf_fp32).astype(np.complex128), workers)
        psi_kp1 = euler_step(psi_k, p2_k, p3_k, c2, k2, p)
        if sample and n % p.t_sample == 0:
            res = residual_calc(psi_kp1, psi_k, p2_k, p3_k, c2, k2, p)
            samples.append(Sample(n, "SSE", rmse_from_residual(res, p, workers)))
        psi = np.real(scaling_step(_ifftn(psi_kp1, workers) * M2, M2))   # copy_matrix: Re
        psi_k = _fftn(psi.astype(np.complex128), workers)
        if sample and (n + 1) % p.t_sample == 0:
            F, fm = free_energy(psi, p, c2, workers)
            samples.append(Sample(n + 1, "FE", F, fm, mean(psi)))
    if sample:
        F, fm = free_energy(psi, p, c2, workers)
        samples.append(Sample(nsteps, "FE", F, fm, mean(psi)))
    return psi, samples


def _full_from_half(res_half: np.ndarray, n: Sequence[int], workers: int) -> np.ndarray:
    """Rebuild a full Hermitian spectrum from its real-transform half (only the
    Re(IFFT) of it is used downstream, sse_calculation.cu:57 / :157)."""
    r = _irfftn(res_half, tuple(n), workers)
    return _fftn(r.astype(np.complex128), workers)


# ----------------------------------------------------------------------------
# initial conditions (initialize.cu:95-202) and the seeded noise IC
# ----------------------------------------------------------------------------
PHILOX_M0 = 0xD2511F53
PHILOX_M1 = 0xCD9E8D57
PHILOX_W0 = 0x9E3779B9
PHILOX_W1 = 0xBB67AE85
_U32 = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr: np.ndarray, key: Tuple[int, int]) -> np.ndarray:
    """Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11 "Parallel random
    numbers: as easy as 1, 2, 3").  ctr: uint64 array (..., 4) holding 32-bit
    words; returns the same shape.  The randn of randn.cu:4-12 is an unseeded
    mt19937 per call (and dead code); a counter-based generator keyed by the
    seed replaces it so that every cell is reproducible."""
    c = np.array(ctr, dtype=np.uint64) & _U32
    k0 = np.uint64(key[0] & 0xFFFFFFFF)
    k1 = np.uint64(key[1] & 0xFFFFFFFF)
    m0 = np.uint64(PHILOX_M0)
    m1 = np.uint64(PHILOX_M1)
    for r in range(10):
        if r > 0:
            k0 = (k0 + np.uint64(PHILOX_W0)) & _U32
            k1 = (k1 + np.uint64(PHILOX_W1)) & _U32
        p0 = m0 * c[..., 0]
        p1 = m1 * c[..., 2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & _U32
        hi1, lo1 = p1 >> np.uint64(32), p1 & _U32
        c = np.stack([(hi1 ^ c[..., 1] ^ k0) & _U32, lo1,
                      (hi0 ^ c[..., 3] ^ k1) & _U32, lo0], axis=-1)
    return c


def _philox_words(c0, c1, key):
    """philox4x32_10 on counters (c0, c1, 0, 0), one 32-bit word per array
    (the same rounds as philox4x32_10, without the (n, 4) stacking); returns
    the first two output words."""
    k0 = np.uint64(key[0] & 0xFFFFFFFF)
    k1 = np.uint64(key[1] & 0xFFFFFFFF)
    m0 = np.uint64(PHILOX_M0)
    m1 = np.uint64(PHILOX_M1)
    s32 = np.uint64(32)
    x0, x1 = c0, c1
    x2 = np.zeros_like(c0)
    x3 = np.zeros_like(c0)
    for r in range(10):
        if r > 0:
            k0 = (k0 + np.uint64(PHILOX_W0)) & _U32
            k1 = (k1 + np.uint64(PHILOX_W1)) & _U32
        p0 = m0 * x0
        p1 = m1 * x2
        x0, x1, x2, x3 = ((p1 >> s32) ^ x1 ^ k0) & _U32, p1 & _U32, ((p0 >> s32) ^ x3 ^ k1) & _U32, p0 & _U32
    return x0, x1


def philox_normals(ncells: int, seed: int, first: int = 0) -> np.ndarray:
    """One N(0,1) per cell c (c = first .. first+ncells-1): Philox4x32-10 with
    counter (c_lo, c_hi, 0, 0) and key (seed_lo, seed_hi); u_i = (r_i + 0.5)
    2^-32; Box-Muller z = sqrt(-2 ln u0) cos(2π u1), in FP64.  Worked in
    chunks of 2^18 cells."""
    key = (seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)
    out = np.empty(ncells, dtype=np.float64)
    chunk = 1 << 18
    for a in range(0, ncells, chunk):
        c = np.arange(first + a, first + min(ncells, a + chunk), dtype=np.uint64)
        r0, r1 = _philox_words(c & _U32, c >> np.uint64(32), key)
        u0 = (r0.astype(np.float64) + 0.5) * (2.0 ** -32)
        u1 = (r1.astype(np.float64) + 0.5) * (2.0 ** -32)
        out[a:a + c.size] = np.sqrt(-2.0 * np.log(u0)) * np.cos(2.0 * PI * u1)
    return out


def noise_ic(p: Params, amplitude: float, seed: int, fix_mean: bool = True) -> np.ndarray:
    """psi = psi0 + amplitude * N(0,1) per cell, then the mean re-fixed to psi0
    as configure_avg_density does (Avg_calc.cu:21-43, initialize.cu:202)."""
    z = philox_normals(p.ncells, seed).reshape(p.n)
    psi = p.psi0 + amplitude * z
    if fix_mean:
        psi = configure_avg_density(psi, p.psi0)
    return psi


def lattice_ic(p: Params, iconfig: int, A_10: float = 0.0, A_11: float = 0.0,
               A_tri: float = 0.0, iseed: int = 0, ishape: int = 0,
               h: float = 0.0, kc: float = 0.0, roa: float = 0.0) -> np.ndarray:
    """2-D initial conditions of initialize.cu:95-202 (square=1, triangular=0,
    liquid=2) with the optional circle/Gaussian seed and the mean fix.
    Coordinates come from linspace_ref (endpoint inclusive, linspace.cu:12)."""
    assert p.dim == 2
    Mx, My = p.n
    Lx, Ly = p.L
    x = linspace_ref(0.0, Lx, Mx)
    y = linspace_ref(0.0, Ly, My)
    X, Y = np.meshgrid(x, y, indexing="ij")
    a_tri = 2.0 / math.sqrt(3.0)
    a_sq = 1.0
    k_sq = 2.0 * PI / a_sq
    k_tri = 2.0 * PI / a_tri
    if iconfig == 1:
        psi = (A_10 * ((3.0 * math.sqrt(2.0) - 2.0) * np.cos(k_sq * Y) - 2.0 * np.cos(k_sq * X))
               - A_11 * np.cos(k_sq * X) * np.cos(k_sq * Y)) + p.psi0
        r = roa * a_sq
    elif iconfig == 0:
        psi = A_tri * (np.cos(k_tri * X) / 2.0
                       + np.cos(k_tri * X / 2.0) * np.cos(math.sqrt(3.0) * k_tri * Y / 2.0)) + p.psi0
        r = roa * a_tri
    elif iconfig == 2:
        psi = np.zeros(p.n) + p.psi0
        r = 0.0
    else:
        raise ValueError("iconfig must be 0, 1 or 2")
    if iseed == 1:
        if ishape == 0:
            d = np.sqrt((X - h) ** 2 + (Y - kc) ** 2)
            psi = np.where(d > r, p.psi0, psi)
        elif ishape == 1:
            g = np.exp(-((X - h) ** 2 + (Y - kc) ** 2) / roa)
            psi = (psi - p.psi0) * g + p.psi0 * (1.0 - g)
    return configure_avg_density(psi, p.psi0)


# ----------------------------------------------------------------------------
# VTK (write_vtk.cu:4-35, read_input.cu:161-241)
# ----------------------------------------------------------------------------
def _fmt6(v: float) -> str:
    """std::ostream default formatting (%g with 6 significant digits)."""
    return "%g" % v


def vtk_text(psi: np.ndarray, Lx: float, Ly: float) -> str:
    """write_vtk.cu:4-35 for a 2-D field indexed psi[i, j] (i < nx): header,
    SPACING Lx/nx Ly/ny 1, then values with x fastest (densityData[j*ny+i])."""
    nx, ny = psi.shape
    lines = ["# vtk DataFile Version 3.0", "Density Data", "ASCII",
             "DATASET STRUCTURED_POINTS", "DIMENSIONS %d %d 1" % (nx, ny),
             "ORIGIN 0 0 0",
             "SPACING %s %s 1" % (_fmt6(float(Lx) / nx), _fmt6(float(Ly) / ny)),
             "POINT_DATA %d" % (nx * ny), "SCALARS DensityData float",
             "LOOKUP_TABLE default"]
    flat = np.real(psi).reshape(-1)
    for i in range(ny):
        for j in range(nx):
            lines.append(_fmt6(float(flat[j * ny + i])))
    return "\n".join(lines) + "\n"
