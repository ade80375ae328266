"""ctypes bindings of libxpfc (include/xpfc.h) and the Python host API."""
from __future__ import annotations

import ctypes as ct
import math
import os
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

XPFC_FP64 = 0
XPFC_FP32_MIXED = 1
XPFC_DCF_SUM = 0
XPFC_DCF_MAX = 1
MAX_PEAKS = 4
# status codes (include/xpfc.h)
XPFC_OK = 0
XPFC_ERR_ARG = 1
XPFC_ERR_IO = 2
XPFC_ERR_PARSE = 3
XPFC_ERR_DEVICE = 4
XPFC_ERR_NOMEM = 5
XPFC_ERR_STATE = 6
XPFC_ERR_NONFINITE = 7

_HERE = os.path.dirname(os.path.abspath(__file__))


def library_path() -> str:
    return os.environ.get("XPFC_LIBRARY", os.path.join(_HERE, "lib", "libxpfc.so"))


class XPFCError(RuntimeError):
    def __init__(self, status: int, where: str, detail: str):
        super().__init__("%s: %s (status %d): %s" % (where, _STATUS.get(status, "?"), status, detail))
        self.status = status


_STATUS = {0: "ok", 1: "invalid argument", 2: "I/O error", 3: "parse error",
           4: "device error", 5: "out of memory", 6: "invalid state", 7: "non-finite result"}


class _Params(ct.Structure):
    _fields_ = [("dim", ct.c_int), ("n", ct.c_int * 3), ("L", ct.c_double * 3),
                ("dt", ct.c_double), ("Mo", ct.c_double), ("eps", ct.c_double),
                ("a", ct.c_double), ("b", ct.c_double), ("sigma", ct.c_double),
                ("psi0", ct.c_double), ("npeaks", ct.c_int),
                ("peak_k", ct.c_double * MAX_PEAKS), ("peak_alpha", ct.c_double * MAX_PEAKS),
                ("peak_rho_beta", ct.c_double * MAX_PEAKS), ("dcf_mode", ct.c_int),
                ("precision", ct.c_int), ("t_sample", ct.c_int), ("int_Mo_residual", ct.c_int),
                ("no_val", ct.c_int), ("t_final", ct.c_int), ("iconfig", ct.c_int),
                ("A_10", ct.c_double), ("A_11", ct.c_double), ("A_tri", ct.c_double),
                ("iseed", ct.c_int), ("ishape", ct.c_int),
                ("h", ct.c_double), ("k", ct.c_double), ("roa", ct.c_double),
                ("device", ct.c_int), ("restart_order", ct.c_int), ("powf_fp32", ct.c_int)]


class _Sample(ct.Structure):
    _fields_ = [("step", ct.c_longlong), ("has_fe", ct.c_int), ("has_sse", ct.c_int),
                ("F_ref", ct.c_double), ("f_mean", ct.c_double), ("mean", ct.c_double),
                ("rmse", ct.c_double)]


class _ServiceClass(ct.Structure):
    _fields_ = [("dim", ct.c_int), ("n", ct.c_int * 3), ("L", ct.c_double * 3), ("slots", ct.c_int),
                ("lane", ct.c_int)]


class _ServiceConfig(ct.Structure):
    _fields_ = [("precision", ct.c_int), ("device", ct.c_int), ("budget_bytes", ct.c_double),
                ("block_steps", ct.c_int), ("lookahead", ct.c_int), ("keep_fields", ct.c_int),
                ("nclasses", ct.c_int), ("classes", ct.POINTER(_ServiceClass))]


class _Job(ct.Structure):
    _fields_ = [("id", ct.c_longlong), ("params", _Params), ("steps", ct.c_longlong),
                ("amplitude", ct.c_double), ("seed", ct.c_ulonglong), ("arrival", ct.c_double)]


class _JobResult(ct.Structure):
    _fields_ = [("id", ct.c_longlong), ("status", ct.c_int), ("steps_done", ct.c_longlong),
                ("F_ref", ct.c_double), ("f_mean", ct.c_double), ("mean", ct.c_double),
                ("t_arrival", ct.c_double), ("t_admitted", ct.c_double), ("t_finished", ct.c_double)]


class _ServiceStats(ct.Structure):
    _fields_ = [(k, ct.c_longlong) for k in ("submitted", "rejected", "done", "nonfinite", "queued",
                                             "running", "blocks", "queue_max")] + \
               [(k, ct.c_double) for k in ("resident_bytes", "peak_resident_bytes", "cell_updates",
                                           "elapsed")]


_lib = None


def load_library(path: Optional[str] = None):
    """Load libxpfc.so; raises if it is absent (there is no fallback path)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or library_path()
    if not os.path.exists(p):
        raise ImportError("libxpfc.so not built at %s; run __graft_entry__.build()" % p)
    L = ct.CDLL(p)
    vp, i, d, ll, sz = ct.c_void_p, ct.c_int, ct.c_double, ct.c_longlong, ct.c_size_t
    P = ct.POINTER
    sig = {
        "xpfc_params_default": (i, [P(_Params)]),
        "xpfc_params_size": (sz, []),
        "xpfc_read_input": (i, [ct.c_char_p, P(_Params)]),
        "xpfc_reference_steps": (ll, [P(_Params)]),
        "xpfc_create": (i, [P(_Params), P(vp)]),
        "xpfc_destroy": (None, [vp]),
        "xpfc_get_params": (i, [vp, P(_Params)]),
        "xpfc_initialize": (i, [vp]),
        "xpfc_initialize_noise": (i, [vp, d, ct.c_ulonglong]),
        "xpfc_randn": (i, [ct.c_ulonglong, ct.c_ulonglong, ll, P(d)]),
        "xpfc_philox4x32_10": (None, [P(ct.c_uint32), P(ct.c_uint32), P(ct.c_uint32)]),
        "xpfc_set_field": (i, [vp, P(d)]),
        "xpfc_get_field": (i, [vp, P(d)]),
        "xpfc_num_cells": (ll, [vp]),
        "xpfc_step": (i, [vp, i]),
        "xpfc_current_step": (ll, [vp]),
        "xpfc_num_samples": (i, [vp]),
        "xpfc_get_samples": (i, [vp, P(_Sample), i]),
        "xpfc_clear_samples": (None, [vp]),
        "xpfc_time_steps": (i, [vp, i, P(ct.c_float)]),
        "xpfc_sync": (i, [vp]),
        "xpfc_time_passes": (i, [vp, i, P(d)]),
        "xpfc_copy_bandwidth": (i, [i, ct.c_size_t, i, P(d)]),
        "xpfc_autotune": (i, [vp, P(i)]),
        "xpfc_set_variants": (i, [vp, P(i)]),
        "xpfc_get_variants": (i, [vp, P(i)]),
        "xpfc_set_graphs": (i, [vp, i]),
        "xpfc_fe_calc": (i, [vp, P(d), P(d)]),
        "xpfc_avg_calc": (i, [vp, P(d)]),
        "xpfc_sse_calc": (i, [vp, P(d)]),
        "xpfc_malloc": (i, [vp, P(vp), sz]),
        "xpfc_free": (i, [vp, vp]),
        "xpfc_memcpy_h2d": (i, [vp, vp, vp, sz]),
        "xpfc_memcpy_d2h": (i, [vp, vp, vp, sz]),
        "xpfc_real_bytes": (sz, [vp]),
        "xpfc_spectrum_bytes": (sz, [vp]),
        "xpfc_power_functions": (i, [vp, i, vp, vp]),
        "xpfc_convolve_fields": (i, [vp, vp, vp]),
        "xpfc_euler_step": (i, [vp, vp, vp, vp, vp]),
        "xpfc_residual_calc": (i, [vp, vp, vp, vp, vp, vp]),
        "xpfc_scaling_step": (i, [vp, vp, i]),
        "xpfc_fft_r2c": (i, [vp, vp, vp]),
        "xpfc_fft_c2r": (i, [vp, vp, vp]),
        "xpfc_write_vtk": (i, [vp, ct.c_char_p]),
        "xpfc_read_vtk": (i, [ct.c_char_p, P(i), P(d), P(d), ll]),
        "xpfc_inst_initialize_vtk": (i, [vp, i, ct.c_char_p]),
        "xpfc_inst_save_checkpoint": (i, [vp, i, ct.c_char_p]),
        "xpfc_inst_load_checkpoint": (i, [vp, i, ct.c_char_p]),
        "xpfc_save_checkpoint": (i, [vp, ct.c_char_p]),
        "xpfc_load_checkpoint": (i, [vp, ct.c_char_p]),
        "xpfc_create_batch": (i, [P(_Params), i, P(vp)]),
        "xpfc_create_mixed": (i, [P(_Params), i, P(vp)]),
        "xpfc_step_budget": (i, [vp, i, P(i)]),
        "xpfc_service_struct_size": (sz, [i]),
        "xpfc_service_create": (i, [P(_ServiceConfig), P(vp)]),
        "xpfc_service_destroy": (None, [vp]),
        "xpfc_service_submit": (i, [vp, P(_Job)]),
        "xpfc_service_poll": (i, [vp]),
        "xpfc_service_idle": (i, [vp]),
        "xpfc_service_run": (i, [vp, P(_Job), i, i]),
        "xpfc_service_results": (i, [vp, P(_JobResult), i]),
        "xpfc_service_field": (ll, [vp, ll, P(d), ll]),
        "xpfc_service_get_stats": (i, [vp, P(_ServiceStats)]),
        "xpfc_get_inst_params": (i, [vp, i, P(_Params)]),
        "xpfc_num_instances": (i, [vp]),
        "xpfc_instance_bytes": (sz, [P(_Params)]),
        "xpfc_query": (i, [vp]),
        "xpfc_probe_finite": (i, [vp, P(i)]),
        "xpfc_set_active": (i, [vp, P(i), i]),
        "xpfc_inst_set_params": (i, [vp, i, P(_Params)]),
        "xpfc_inst_reset": (i, [vp, i]),
        "xpfc_inst_initialize": (i, [vp, i]),
        "xpfc_inst_initialize_noise": (i, [vp, i, d, ct.c_ulonglong]),
        "xpfc_inst_initialize_noise_noalloc": (i, [vp, i, d, ct.c_ulonglong]),
        "xpfc_inst_set_field": (i, [vp, i, P(d)]),
        "xpfc_inst_get_field": (i, [vp, i, P(d)]),
        "xpfc_inst_current_step": (ll, [vp, i]),
        "xpfc_inst_num_samples": (i, [vp, i]),
        "xpfc_inst_get_samples": (i, [vp, i, P(_Sample), i]),
        "xpfc_inst_clear_samples": (None, [vp, i]),
        "xpfc_inst_fe_calc": (i, [vp, i, P(d), P(d)]),
        "xpfc_inst_avg_calc": (i, [vp, i, P(d)]),
        "xpfc_inst_sse_calc": (i, [vp, i, P(d)]),
        "xpfc_inst_write_vtk": (i, [vp, i, ct.c_char_p]),
        "xpfc_strerror": (ct.c_char_p, [i]),
        "xpfc_last_error": (ct.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    if L.xpfc_params_size() != ct.sizeof(_Params):
        raise ImportError("xpfc_params layout mismatch: library %d bytes, binding %d"
                          % (L.xpfc_params_size(), ct.sizeof(_Params)))
    for which, st in enumerate((_ServiceClass, _ServiceConfig, _Job, _JobResult, _ServiceStats)):
        if L.xpfc_service_struct_size(which) != ct.sizeof(st):
            raise ImportError("%s layout mismatch: library %d bytes, binding %d"
                              % (st.__name__, L.xpfc_service_struct_size(which), ct.sizeof(st)))
    _lib = L
    return L


def lib():
    return load_library()


def _check(status: int, where: str):
    if status != 0:
        detail = lib().xpfc_last_error().decode(errors="replace")
        raise XPFCError(status, where, detail)


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


# ---------------------------------------------------------------------------
@dataclass
class Params:
    """xpfc_params (read_input.cu:19-115, globals.cuh:40-52)."""
    n: Tuple[int, ...] = (256, 256)
    L: Tuple[float, ...] = (32.0, 32.0)
    dt: float = 0.5
    Mo: float = 1.0
    eps: float = 0.0
    a: float = 1.0
    b: float = 1.0 / 3.0
    sigma: float = 0.1
    psi0: float = 0.05
    peaks: Sequence[Tuple[float, float, float]] = (
        (2.0 * math.pi, 1.0, 4.0), (2.0 * math.pi * math.sqrt(2.0), 1.0, 4.0 / math.sqrt(2.0)))
    dcf_mode: int = XPFC_DCF_SUM
    precision: int = XPFC_FP64
    t_sample: int = 100
    int_Mo_residual: int = 1
    no_val: int = 0
    t_final: int = 0
    iconfig: int = 2
    A_10: float = 0.0
    A_11: float = 0.0
    A_tri: float = 0.0
    iseed: int = 0
    ishape: int = 0
    h: float = 0.0
    k: float = 0.0
    roa: float = 0.0
    device: int = 0
    restart_order: int = 0
    powf_fp32: int = 0

    @property
    def dim(self) -> int:
        return len(self.n)

    def to_c(self) -> _Params:
        c = _Params()
        _check(lib().xpfc_params_default(ct.byref(c)), "xpfc_params_default")
        c.dim = len(self.n)
        for i in range(3):
            c.n[i] = int(self.n[i]) if i < len(self.n) else 1
            c.L[i] = float(self.L[i]) if i < len(self.L) else 1.0
        for f in ("dt", "Mo", "eps", "a", "b", "sigma", "psi0", "A_10", "A_11", "A_tri",
                  "h", "k", "roa"):
            setattr(c, f, float(getattr(self, f)))
        for f in ("dcf_mode", "precision", "t_sample", "int_Mo_residual", "no_val", "t_final",
                  "iconfig", "iseed", "ishape", "device", "restart_order", "powf_fp32"):
            setattr(c, f, int(getattr(self, f)))
        if not 1 <= len(self.peaks) <= MAX_PEAKS:
            raise ValueError("1..4 DCF peaks")
        c.npeaks = len(self.peaks)
        for i, (kk, al, rb) in enumerate(self.peaks):
            c.peak_k[i], c.peak_alpha[i], c.peak_rho_beta[i] = kk, al, rb
        return c

    @classmethod
    def from_c(cls, c: _Params) -> "Params":
        nd = c.dim
        return cls(n=tuple(c.n[i] for i in range(nd)), L=tuple(c.L[i] for i in range(nd)),
                   dt=c.dt, Mo=c.Mo, eps=c.eps, a=c.a, b=c.b, sigma=c.sigma, psi0=c.psi0,
                   peaks=tuple((c.peak_k[i], c.peak_alpha[i], c.peak_rho_beta[i])
                               for i in range(c.npeaks)),
                   dcf_mode=c.dcf_mode, precision=c.precision, t_sample=c.t_sample,
                   int_Mo_residual=c.int_Mo_residual, no_val=c.no_val, t_final=c.t_final,
                   iconfig=c.iconfig, A_10=c.A_10, A_11=c.A_11, A_tri=c.A_tri, iseed=c.iseed,
                   ishape=c.ishape, h=c.h, k=c.k, roa=c.roa, device=c.device,
                   restart_order=c.restart_order, powf_fp32=c.powf_fp32)


@dataclass
class Sample:
    step: int
    has_fe: bool
    has_sse: bool
    F_ref: float
    f_mean: float
    mean: float
    rmse: float

    def line(self) -> str:
        """The reference's printed line (sum_reduction.cu:64,77)."""
        s = ""
        if self.has_fe:
            s += "iteration: %d F = %0.10f" % (self.step, self.F_ref)
        if self.has_sse:
            s += " SSE_gpu = %0.5e \n" % self.rmse
        return s


def read_input(path: str) -> Params:
    """read_input (read_input.cu:8-121)."""
    c = _Params()
    _check(lib().xpfc_read_input(path.encode(), ct.byref(c)), "xpfc_read_input")
    return Params.from_c(c)


def reference_steps(p: Params) -> int:
    c = p.to_c()
    return int(lib().xpfc_reference_steps(ct.byref(c)))


def read_vtk(path: str):
    """Legacy VTK structured points -> (dims, spacing, values in file order)."""
    dims = (ct.c_int * 3)()
    sp = (ct.c_double * 3)()
    _check(lib().xpfc_read_vtk(path.encode(), dims, sp, None, 0), "xpfc_read_vtk")
    n = dims[0] * dims[1] * dims[2]
    out = np.empty(n, dtype=np.float64)
    _check(lib().xpfc_read_vtk(path.encode(), dims, sp, _dptr(out), n), "xpfc_read_vtk")
    return tuple(dims), tuple(sp), out


def instance_bytes(p: Params) -> int:
    """Device bytes one instance of this grid and precision occupies."""
    c = p.to_c()
    return int(lib().xpfc_instance_bytes(ct.byref(c)))


def copy_bandwidth(nbytes: int, reps: int = 20, device: int = 0) -> float:
    """GB/s (read + written) of a plain streaming copy of `nbytes` on the GPU:
    the practical HBM ceiling the passes are compared with."""
    g = ct.c_double()
    _check(lib().xpfc_copy_bandwidth(int(device), int(nbytes), int(reps), ct.byref(g)), "xpfc_copy_bandwidth")
    return g.value


def randn(seed: int, count: int, first: int = 0) -> np.ndarray:
    """Seeded replacement of randn() (randn.cu:4-12)."""
    out = np.empty(count, dtype=np.float64)
    _check(lib().xpfc_randn(seed, first, count, _dptr(out)), "xpfc_randn")
    return out


def philox4x32_10(ctr: Sequence[int], key: Sequence[int]) -> List[int]:
    c = (ct.c_uint32 * 4)(*ctr)
    k = (ct.c_uint32 * 2)(*key)
    o = (ct.c_uint32 * 4)()
    lib().xpfc_philox4x32_10(c, k, o)
    return list(o)


class DeviceBuffer:
    """A device allocation owned by a context (operator-level API)."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = ct.c_void_p()
        _check(lib().xpfc_malloc(ctx.handle, ct.byref(p), self.nbytes), "xpfc_malloc")
        self.ptr = p

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes, (a.nbytes, self.nbytes)
        _check(lib().xpfc_memcpy_h2d(self.ctx.handle, self.ptr, a.ctypes.data_as(ct.c_void_p),
                                     self.nbytes), "xpfc_memcpy_h2d")

    def download(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes == self.nbytes
        _check(lib().xpfc_memcpy_d2h(self.ctx.handle, out.ctypes.data_as(ct.c_void_p), self.ptr,
                                     self.nbytes), "xpfc_memcpy_d2h")
        return out

    def free(self):
        if self.ptr:
            _check(lib().xpfc_free(self.ctx.handle, self.ptr), "xpfc_free")
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.ctx.handle:
                lib().xpfc_free(self.ctx.handle, self.ptr)
        except Exception:
            pass


class Context:
    """An xpfc_ctx: device fields, tables and stream for one simulation."""

    def __init__(self, params: Params):
        self.params = params
        h = ct.c_void_p()
        c = params.to_c()
        _check(lib().xpfc_create(ct.byref(c), ct.byref(h)), "xpfc_create")
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            lib().xpfc_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # dtype of device fields
    @property
    def real_dtype(self):
        return np.float32 if self.params.precision == XPFC_FP32_MIXED else np.float64

    @property
    def complex_dtype(self):
        return np.complex64 if self.params.precision == XPFC_FP32_MIXED else np.complex128

    @property
    def shape(self):
        return tuple(self.params.n)

    @property
    def spectrum_shape(self):
        n = self.params.n
        return tuple(n[:-1]) + (n[-1] // 2 + 1,)

    # --- initial conditions / state
    def initialize(self):
        """initialize() (initialize.cu:6-238)."""
        _check(lib().xpfc_initialize(self.handle), "xpfc_initialize")

    def initialize_noise(self, amplitude: float, seed: int):
        _check(lib().xpfc_initialize_noise(self.handle, float(amplitude), int(seed)),
               "xpfc_initialize_noise")

    def set_field(self, psi: np.ndarray):
        a = np.ascontiguousarray(psi, dtype=np.float64)
        assert a.shape == self.shape, (a.shape, self.shape)
        _check(lib().xpfc_set_field(self.handle, _dptr(a)), "xpfc_set_field")

    def get_field(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64)
        _check(lib().xpfc_get_field(self.handle, _dptr(out)), "xpfc_get_field")
        return out

    # --- hot path
    def step(self, nsteps: int):
        _check(lib().xpfc_step(self.handle, int(nsteps)), "xpfc_step")

    def time_steps(self, nsteps: int) -> float:
        ms = ct.c_float()
        _check(lib().xpfc_time_steps(self.handle, int(nsteps), ct.byref(ms)), "xpfc_time_steps")
        return float(ms.value)

    def time_passes(self, nsteps: int):
        """Average ms of [axis-0 fused, axis-1 inverse, real-space, axis-1 forward, whole step]."""
        ms = (ct.c_double * 5)()
        _check(lib().xpfc_time_passes(self.handle, int(nsteps), ms), "xpfc_time_passes")
        return list(ms)

    def autotune(self) -> Tuple[int, int, int]:
        v = (ct.c_int * 3)()
        _check(lib().xpfc_autotune(self.handle, v), "xpfc_autotune")
        return tuple(v)

    def set_variants(self, tc_x: int = 0, tc_y: int = 0, ntz: int = 0):
        """Pin the pass variants (0 = keep): axis-0 tile width (4/8/16); axis-1
        tile width (+1: workgroups dealt to the XCDs in pairs); real-space
        workgroup size 64..512 (+1: no occupancy hint)."""
        v = (ct.c_int * 3)(tc_x, tc_y, ntz)
        _check(lib().xpfc_set_variants(self.handle, v), "xpfc_set_variants")

    def variants(self) -> Tuple[int, int, int]:
        v = (ct.c_int * 3)()
        _check(lib().xpfc_get_variants(self.handle, v), "xpfc_get_variants")
        return tuple(v)

    def set_graphs(self, on: bool):
        _check(lib().xpfc_set_graphs(self.handle, int(bool(on))), "xpfc_set_graphs")

    def sync(self):
        _check(lib().xpfc_sync(self.handle), "xpfc_sync")

    @property
    def current_step(self) -> int:
        return int(lib().xpfc_current_step(self.handle))

    def samples(self) -> List[Sample]:
        n = lib().xpfc_num_samples(self.handle)
        if n < 0:
            _check(-n, "xpfc_num_samples")
        arr = (_Sample * max(n, 1))()
        got = lib().xpfc_get_samples(self.handle, arr, n)
        if got < 0:
            _check(-got, "xpfc_get_samples")
        return [Sample(int(s.step), bool(s.has_fe), bool(s.has_sse), s.F_ref, s.f_mean, s.mean,
                       s.rmse) for s in arr[:got]]

    def clear_samples(self):
        lib().xpfc_clear_samples(self.handle)

    # --- diagnostics
    def fe_calc(self) -> Tuple[float, float]:
        """FE_calc (FE_calc.cu:3-68): (F_ref, f_mean)."""
        F, fm = ct.c_double(), ct.c_double()
        _check(lib().xpfc_fe_calc(self.handle, ct.byref(F), ct.byref(fm)), "xpfc_fe_calc")
        return F.value, fm.value

    def avg_calc(self) -> float:
        """mean (Avg_calc.cu:3-18)."""
        m = ct.c_double()
        _check(lib().xpfc_avg_calc(self.handle, ct.byref(m)), "xpfc_avg_calc")
        return m.value

    def sse_calc(self) -> float:
        """SSE_calc_wrapper (sse_calculation.cu:3-116): RMSE of the last sampled update."""
        r = ct.c_double()
        _check(lib().xpfc_sse_calc(self.handle, ct.byref(r)), "xpfc_sse_calc")
        return r.value

    def write_vtk(self, path: str):
        """writeVTKFile (write_vtk.cu:4-35)."""
        _check(lib().xpfc_write_vtk(self.handle, path.encode()), "xpfc_write_vtk")

    def initialize_vtk(self, path: str):
        """read_density_field (read_input.cu:161-241) + the average-density fix."""
        _check(lib().xpfc_inst_initialize_vtk(self.handle, 0, path.encode()),
               "xpfc_inst_initialize_vtk")

    def save_checkpoint(self, path: str):
        _check(lib().xpfc_save_checkpoint(self.handle, path.encode()), "xpfc_save_checkpoint")

    def load_checkpoint(self, path: str):
        _check(lib().xpfc_load_checkpoint(self.handle, path.encode()), "xpfc_load_checkpoint")

    # --- operator level (host arrays in, host arrays out; device buffers inside)
    def _buf(self, a: np.ndarray) -> DeviceBuffer:
        b = DeviceBuffer(self, a.nbytes)
        b.upload(a)
        return b

    def _real_in(self, a):
        a = np.ascontiguousarray(a, dtype=self.real_dtype)
        assert a.shape == self.shape
        return self._buf(a)

    def _spec_in(self, a):
        a = np.ascontiguousarray(a, dtype=self.complex_dtype)
        assert a.shape == self.spectrum_shape, (a.shape, self.spectrum_shape)
        return self._buf(a)

    def power_functions(self, psi: np.ndarray, p: int) -> np.ndarray:
        """power_function_wrapper (power_functions.cu:4-53)."""
        s = self._real_in(psi)
        d = DeviceBuffer(self, s.nbytes)
        _check(lib().xpfc_power_functions(self.handle, int(p), s.ptr, d.ptr), "xpfc_power_functions")
        return d.download(self.real_dtype, self.shape)

    def convolve_fields(self, psi_k: np.ndarray) -> np.ndarray:
        """convolution_wrapper (convolve_fields.cu:3-51)."""
        s = self._spec_in(psi_k)
        d = DeviceBuffer(self, s.nbytes)
        _check(lib().xpfc_convolve_fields(self.handle, s.ptr, d.ptr), "xpfc_convolve_fields")
        return d.download(self.complex_dtype, self.spectrum_shape)

    def euler_step(self, psi_k, p2_k, p3_k) -> np.ndarray:
        """euler_step_wrapper (euler_step.cu:3-63)."""
        a, b, c = self._spec_in(psi_k), self._spec_in(p2_k), self._spec_in(p3_k)
        d = DeviceBuffer(self, a.nbytes)
        _check(lib().xpfc_euler_step(self.handle, a.ptr, b.ptr, c.ptr, d.ptr), "xpfc_euler_step")
        return d.download(self.complex_dtype, self.spectrum_shape)

    def residual_calc(self, psi_kp1, psi_k, p2_k, p3_k) -> np.ndarray:
        """residual_calc (sse_calculation.cu:120-147)."""
        bufs = [self._spec_in(x) for x in (psi_kp1, psi_k, p2_k, p3_k)]
        d = DeviceBuffer(self, bufs[0].nbytes)
        _check(lib().xpfc_residual_calc(self.handle, *[b.ptr for b in bufs], d.ptr),
               "xpfc_residual_calc")
        return d.download(self.complex_dtype, self.spectrum_shape)

    def scaling_step(self, f: np.ndarray) -> np.ndarray:
        """scaling_step_wrapper (scaling_step.cu:4-53)."""
        spectral = np.iscomplexobj(f)
        b = self._spec_in(f) if spectral else self._real_in(f)
        _check(lib().xpfc_scaling_step(self.handle, b.ptr, int(spectral)), "xpfc_scaling_step")
        if spectral:
            return b.download(self.complex_dtype, self.spectrum_shape)
        return b.download(self.real_dtype, self.shape)

    def fft_r2c(self, real: np.ndarray) -> np.ndarray:
        s = self._real_in(real)
        d = DeviceBuffer(self, int(np.prod(self.spectrum_shape)) * np.dtype(self.complex_dtype).itemsize)
        _check(lib().xpfc_fft_r2c(self.handle, s.ptr, d.ptr), "xpfc_fft_r2c")
        return d.download(self.complex_dtype, self.spectrum_shape)

    def fft_c2r(self, spec: np.ndarray) -> np.ndarray:
        s = self._spec_in(spec)
        d = DeviceBuffer(self, int(np.prod(self.shape)) * np.dtype(self.real_dtype).itemsize)
        _check(lib().xpfc_fft_c2r(self.handle, s.ptr, d.ptr), "xpfc_fft_c2r")
        return d.download(self.real_dtype, self.shape)


class BatchContext:
    """Several simulations of one grid in one context (xpfc_create_batch).

    Every instance has its own physics, field, step counter and sample log;
    step() advances the active instances with one launch per pass.  No
    counterpart in the reference (one simulation per process, main.cu:20-435).
    """

    def __init__(self, params: Sequence[Params]):
        if not params:
            raise ValueError("at least one instance")
        self.params = list(params)
        arr = (_Params * len(self.params))(*[q.to_c() for q in self.params])
        h = ct.c_void_p()
        _check(lib().xpfc_create_batch(arr, len(self.params), ct.byref(h)), "xpfc_create_batch")
        self.handle = h

    def query(self) -> bool:
        """True when every queued step has finished (never blocks)."""
        r = lib().xpfc_query(self.handle)
        if r < 0:
            _check(-r, "xpfc_query")
        return r == 1

    def probe_finite(self) -> List[bool]:
        """Per instance: False once its field has diverged (waits for queued steps)."""
        v = (ct.c_int * self.ninst)()
        _check(lib().xpfc_probe_finite(self.handle, v), "xpfc_probe_finite")
        return [bool(x) for x in v]

    close = Context.close
    __del__ = Context.__del__
    __enter__ = Context.__enter__
    __exit__ = Context.__exit__
    step = Context.step
    time_steps = Context.time_steps
    time_passes = Context.time_passes
    autotune = Context.autotune
    set_variants = Context.set_variants
    variants = Context.variants
    set_graphs = Context.set_graphs
    sync = Context.sync

    @property
    def ninst(self) -> int:
        return int(lib().xpfc_num_instances(self.handle))

    @property
    def shape(self):
        return tuple(self.params[0].n)

    def set_active(self, inst: Optional[Sequence[int]] = None):
        if inst is None:
            _check(lib().xpfc_set_active(self.handle, None, 0), "xpfc_set_active")
        else:
            a = (ct.c_int * max(len(inst), 1))(*inst)
            _check(lib().xpfc_set_active(self.handle, a, len(inst)), "xpfc_set_active")

    def set_params(self, s: int, p: Params):
        c = p.to_c()
        _check(lib().xpfc_inst_set_params(self.handle, s, ct.byref(c)), "xpfc_inst_set_params")
        self.params[s] = p

    def reset(self, s: int):
        _check(lib().xpfc_inst_reset(self.handle, s), "xpfc_inst_reset")

    def initialize(self, s: int):
        _check(lib().xpfc_inst_initialize(self.handle, s), "xpfc_inst_initialize")

    def initialize_noise(self, s: int, amplitude: float, seed: int):
        _check(lib().xpfc_inst_initialize_noise(self.handle, s, float(amplitude), int(seed)),
               "xpfc_inst_initialize_noise")

    def initialize_noise_noalloc(self, s: int, amplitude: float, seed: int):
        """initialize_noise with no device or host allocation (same field)."""
        _check(lib().xpfc_inst_initialize_noise_noalloc(self.handle, s, float(amplitude), int(seed)),
               "xpfc_inst_initialize_noise_noalloc")

    def shape_of(self, s: int):
        return self.shape

    def set_field(self, s: int, psi: np.ndarray):
        a = np.ascontiguousarray(psi, dtype=np.float64)
        assert a.shape == self.shape_of(s), (a.shape, self.shape_of(s))
        _check(lib().xpfc_inst_set_field(self.handle, s, _dptr(a)), "xpfc_inst_set_field")

    def get_field(self, s: int) -> np.ndarray:
        out = np.empty(self.shape_of(s), dtype=np.float64)
        _check(lib().xpfc_inst_get_field(self.handle, s, _dptr(out)), "xpfc_inst_get_field")
        return out

    def current_step(self, s: int) -> int:
        return int(lib().xpfc_inst_current_step(self.handle, s))

    def samples(self, s: int) -> List[Sample]:
        n = lib().xpfc_inst_num_samples(self.handle, s)
        if n < 0:
            _check(-n, "xpfc_inst_num_samples")
        arr = (_Sample * max(n, 1))()
        got = lib().xpfc_inst_get_samples(self.handle, s, arr, n)
        if got < 0:
            _check(-got, "xpfc_inst_get_samples")
        return [Sample(int(r.step), bool(r.has_fe), bool(r.has_sse), r.F_ref, r.f_mean, r.mean,
                       r.rmse) for r in arr[:got]]

    def clear_samples(self, s: int):
        lib().xpfc_inst_clear_samples(self.handle, s)

    def fe_calc(self, s: int) -> Tuple[float, float]:
        F, fm = ct.c_double(), ct.c_double()
        _check(lib().xpfc_inst_fe_calc(self.handle, s, ct.byref(F), ct.byref(fm)),
               "xpfc_inst_fe_calc")
        return F.value, fm.value

    def avg_calc(self, s: int) -> float:
        m = ct.c_double()
        _check(lib().xpfc_inst_avg_calc(self.handle, s, ct.byref(m)), "xpfc_inst_avg_calc")
        return m.value

    def sse_calc(self, s: int) -> float:
        r = ct.c_double()
        _check(lib().xpfc_inst_sse_calc(self.handle, s, ct.byref(r)), "xpfc_inst_sse_calc")
        return r.value

    def write_vtk(self, s: int, path: str):
        _check(lib().xpfc_inst_write_vtk(self.handle, s, path.encode()), "xpfc_inst_write_vtk")

    def initialize_vtk(self, s: int, path: str):
        _check(lib().xpfc_inst_initialize_vtk(self.handle, s, path.encode()),
               "xpfc_inst_initialize_vtk")

    def save_checkpoint(self, s: int, path: str):
        _check(lib().xpfc_inst_save_checkpoint(self.handle, s, path.encode()),
               "xpfc_inst_save_checkpoint")

    def load_checkpoint(self, s: int, path: str):
        _check(lib().xpfc_inst_load_checkpoint(self.handle, s, path.encode()),
               "xpfc_inst_load_checkpoint")


class MixedContext(BatchContext):
    """Simulations of several grids in one context (xpfc_create_mixed): each
    step advances every active instance of every grid with one launch per
    pass, and runs of unsampled steps replay captured graphs of those launches.
    Instance s keeps the grid of params[s]; its field equals the one a
    one-instance context gives, bit for bit.  No counterpart in the reference
    (one simulation per process, main.cu:20-435)."""

    def __init__(self, params: Sequence[Params]):
        if not params:
            raise ValueError("at least one instance")
        self.params = list(params)
        arr = (_Params * len(self.params))(*[q.to_c() for q in self.params])
        h = ct.c_void_p()
        _check(lib().xpfc_create_mixed(arr, len(self.params), ct.byref(h)), "xpfc_create_mixed")
        self.handle = h

    @property
    def shape(self):
        raise AttributeError("a mixed context has one shape per instance: use shape_of(s)")

    def shape_of(self, s: int):
        return tuple(self.params[s].n)

