"""Parameter sets shared by the parity tests (oracle <-> library)."""
import math

from oracle import xpfc_oracle as O
from xpfc_cuda_amd import Params, XPFC_DCF_MAX, XPFC_DCF_SUM, XPFC_FP32_MIXED, XPFC_FP64

TWO_PI = 2.0 * math.pi


def square_peaks(rho_beta=1.0):
    """Square lattice: k1 = 2π, k2 = 2√2π, α = 1, exp(-σ²k²/(2ρβ))."""
    return [(TWO_PI, 1.0, rho_beta), (TWO_PI * math.sqrt(2.0), 1.0, rho_beta)]


def make(n, dx=0.125, sigma=0.1, psi0=0.05, peaks=None, mode="max", prec=XPFC_FP64,
         t_sample=100, dt=0.5, eps=0.0, Mo=1.0, a=1.0, b=1.0 / 3.0, int_Mo=1, powf_fp32=0):
    """Return (oracle Params, library Params) for the same physics."""
    peaks = list(peaks if peaks is not None else square_peaks())
    L = tuple(float(m) * dx for m in n)
    op = O.Params(n=tuple(n), L=L, dt=dt, M=Mo, eps=eps, a=a, b=b, sigma=sigma,
                  peaks=[O.Peak(k, al, rb) for (k, al, rb) in peaks], dcf_mode=mode,
                  psi0=psi0, t_sample=t_sample, powf_fp32=bool(powf_fp32), int_Mo=bool(int_Mo))
    lp = Params(n=tuple(n), L=L, dt=dt, Mo=Mo, eps=eps, a=a, b=b, sigma=sigma, psi0=psi0,
                peaks=tuple(peaks), dcf_mode=XPFC_DCF_MAX if mode == "max" else XPFC_DCF_SUM,
                precision=prec, t_sample=t_sample, int_Mo_residual=int_Mo, powf_fp32=powf_fp32)
    return op, lp


# crystallising case: Ĉ2 peak (0.908) above the liquid threshold 1-2a n0+3b n0² (0.9025)
def growth_case(n, prec=XPFC_FP64, **kw):
    return make(n, sigma=0.07, prec=prec, **kw)
