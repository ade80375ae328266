"""FP64 real-to-complex transform by rocFFT through ctypes, as an independent
second FFT implementation for the GPU tests (test-only; the library never
calls rocFFT).  The output is numpy.fft.rfftn's layout: the half spectrum of a
C-ordered array, last axis n/2 + 1."""
import ctypes as ct

import numpy as np

_HIP_H2D, _HIP_D2H = 1, 2
_PLACE_NOTINPLACE = 1
_REAL_FORWARD = 2
_DOUBLE = 1

_libs = {}


def _load():
    if not _libs:
        hip = ct.CDLL("/opt/rocm/lib/libamdhip64.so")
        fft = ct.CDLL("/opt/rocm/lib/librocfft.so")
        hip.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
        hip.hipFree.argtypes = [ct.c_void_p]
        hip.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
        fft.rocfft_plan_create.argtypes = [ct.POINTER(ct.c_void_p), ct.c_int, ct.c_int, ct.c_int,
                                           ct.c_size_t, ct.POINTER(ct.c_size_t), ct.c_size_t,
                                           ct.c_void_p]
        fft.rocfft_plan_get_work_buffer_size.argtypes = [ct.c_void_p, ct.POINTER(ct.c_size_t)]
        fft.rocfft_execution_info_create.argtypes = [ct.POINTER(ct.c_void_p)]
        fft.rocfft_execution_info_set_work_buffer.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t]
        fft.rocfft_execution_info_destroy.argtypes = [ct.c_void_p]
        fft.rocfft_execute.argtypes = [ct.c_void_p, ct.POINTER(ct.c_void_p),
                                       ct.POINTER(ct.c_void_p), ct.c_void_p]
        fft.rocfft_plan_destroy.argtypes = [ct.c_void_p]
        if fft.rocfft_setup() != 0:
            raise RuntimeError("rocfft_setup failed")
        _libs["hip"], _libs["fft"] = hip, fft
    return _libs["hip"], _libs["fft"]


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed (%d)" % (what, rc))


def rfftn(x: np.ndarray) -> np.ndarray:
    hip, fft = _load()
    x = np.ascontiguousarray(x, dtype=np.float64)
    out_shape = x.shape[:-1] + (x.shape[-1] // 2 + 1,)
    out = np.empty(out_shape, dtype=np.complex128)
    lengths = (ct.c_size_t * x.ndim)(*reversed(x.shape))  # fastest axis first
    din, dout, work = ct.c_void_p(), ct.c_void_p(), ct.c_void_p()
    plan, info = ct.c_void_p(), ct.c_void_p()
    try:
        _ok(hip.hipMalloc(ct.byref(din), x.nbytes), "hipMalloc")
        _ok(hip.hipMalloc(ct.byref(dout), out.nbytes), "hipMalloc")
        _ok(hip.hipMemcpy(din, x.ctypes.data, x.nbytes, _HIP_H2D), "hipMemcpy")
        _ok(fft.rocfft_plan_create(ct.byref(plan), _PLACE_NOTINPLACE, _REAL_FORWARD, _DOUBLE,
                                   x.ndim, lengths, 1, None), "rocfft_plan_create")
        wsize = ct.c_size_t(0)
        _ok(fft.rocfft_plan_get_work_buffer_size(plan, ct.byref(wsize)), "work size")
        _ok(fft.rocfft_execution_info_create(ct.byref(info)), "execution info")
        if wsize.value:
            _ok(hip.hipMalloc(ct.byref(work), wsize.value), "hipMalloc")
            _ok(fft.rocfft_execution_info_set_work_buffer(info, work, wsize.value), "work buffer")
        ins = (ct.c_void_p * 1)(din)
        outs = (ct.c_void_p * 1)(dout)
        _ok(fft.rocfft_execute(plan, ins, outs, info), "rocfft_execute")
        _ok(hip.hipMemcpy(out.ctypes.data, dout, out.nbytes, _HIP_D2H), "hipMemcpy")
    finally:
        if info:
            fft.rocfft_execution_info_destroy(info)
        if plan:
            fft.rocfft_plan_destroy(plan)
        for p in (din, dout, work):
            if p:
                hip.hipFree(p)
    return out
