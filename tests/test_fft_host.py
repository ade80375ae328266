"""The device FFT stages (xpfc_cuda_amd/csrc/fft_core.hpp) run on the CPU.

tests/fft_harness.cpp spreads a line over L/E host threads as a workgroup
does (one std::thread per GPU thread, __syncthreads as a barrier, the LDS a
bounds-checked shared array) and compares each forward and inverse transform
with the exact DFT in long double.  Every compiled line length is covered
(2^k, 3*2^k, 5*2^k: the radix-3/5/6/10/12/20 butterflies and the mixed stage
schedule), with the points per thread of the strided passes and of the
real-space pass in both precisions.  Nothing here needs a GPU.
"""
import ctypes
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def compiled_lengths():
    txt = open(os.path.join(ROOT, "xpfc_cuda_amd", "csrc", "xpfc_lengths.hpp")).read()
    block = txt[txt.index("#define XPFC_LENGTHS(X)"):]
    block = block[:block.index("\n\n")]
    return sorted(int(v) for v in re.findall(r"X\((\d+)\)", block))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("fft") / "libfft_harness.so")
    subprocess.run([gxx, "-std=c++20", "-O1", "-shared", "-fPIC", "-pthread",
                    "-I", os.path.join(ROOT, "tests", "fft_emu"),
                    os.path.join(ROOT, "tests", "fft_harness.cpp"), "-o", out], check=True)
    lib = ctypes.CDLL(out)
    lib.fft_check.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint,
                              ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]
    lib.fft_check.restype = ctypes.c_int
    return lib


def run(lib, L, shape, prec, direction, seed=1, flip=0):
    err = ctypes.c_double(-1.0)
    e = ctypes.c_int(0)
    rc = lib.fft_check(L, shape, prec, direction, seed, flip, ctypes.byref(err), ctypes.byref(e))
    assert rc == 0, "length %d: %s" % (L, {1: "not in the harness", 2: "LDS index out of the line"}.get(rc, rc))
    return err.value, e.value


@pytest.mark.parametrize("L", compiled_lengths())
def test_every_compiled_length_matches_the_exact_dft(harness, L):
    # FP64 within 1e-14 of max|X| (about log2(L) roundings of 1.1e-16), FP32 within 2e-6
    for prec, tol in ((0, 1e-14), (1, 2e-6)):
        for shape in (0, 1):
            for direction in (-1, 1):
                err, e = run(harness, L, shape, prec, direction, seed=L + 7 * shape)
                assert L % e == 0
                assert err <= tol, (L, prec, shape, direction, e, err)


@pytest.mark.parametrize("L", [3, 5, 6, 10])
def test_small_radices_alone(harness, L):
    # one thread, one stage: the radix-3/5/6/10 butterflies themselves
    for direction in (-1, 1):
        err, e = run(harness, L, 0, 0, direction)
        assert e == L
        assert err <= 1e-15, (L, direction, err)


@pytest.mark.parametrize("L", [12, 20, 96, 160])
def test_a_wrong_sign_is_seen(harness, L):
    # the comparison is sharp: the same transforms checked against the DFT of
    # the opposite sign (a sign slip in a butterfly) fail by O(1)
    for direction in (-1, 1):
        err, _ = run(harness, L, 0, 0, direction, flip=1)
        assert err > 0.1, (L, direction, err)


def test_points_per_thread_leave_whole_wavefront_lines(harness):
    # the real-space pass exchanges mirror bins across the lanes of a line:
    # its threads per line must be a power of two of at most 64 for every half
    # line n2/2 of a supported extent (n2 <= 2048)
    for L in [v for v in compiled_lengths() if v <= 1024]:
        for prec in (0, 1):
            _, e = run(harness, L, 1, prec, -1)
            tpl = L // e
            assert tpl & (tpl - 1) == 0 and tpl <= 64, (L, prec, e)
