"""The C ABI library: loads without a GPU, exports every declared symbol, and
its host-only entry points (params, read_input, Philox/randn) behave."""
import math
import os
import re

import numpy as np
import pytest

import xpfc_cuda_amd as X
from oracle import xpfc_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "xpfc.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(xpfc_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = X.load_library()
    names = declared_functions()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_params_default_matches_reference_constants():
    p = X.Params()
    c = p.to_c()
    d = X.api._Params()
    import ctypes
    assert X.lib().xpfc_params_default(ctypes.byref(d)) == 0
    # initialize.cu:10-17
    assert d.npeaks == 2
    assert d.peak_k[0] == 2.0 * math.pi and d.peak_k[1] == 2.0 * math.pi * math.sqrt(2.0)
    assert d.peak_rho_beta[0] == 4.0 and d.peak_rho_beta[1] == 4.0 / math.sqrt(2.0)
    assert d.dcf_mode == X.XPFC_DCF_SUM
    assert c.dim == 2 and c.n[2] == 1


# Random123 known-answer vectors for Philox4x32-10 (Salmon et al., SC'11)
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_philox_kat_library(ctr, key, expect):
    assert tuple(X.philox4x32_10(ctr, key)) == expect


def test_randn_library_matches_oracle():
    seed = 12345678901
    got = X.randn(seed, 4096, first=0)
    ref = O.philox_normals(4096, seed)
    assert np.max(np.abs(got - ref)) <= 4e-15
    # offset draws are the same stream
    assert np.array_equal(X.randn(seed, 10, first=100), got[100:110])


PFC_INPUT = """1          no_val
0.05 0.1   psi_0 s
1.0 3.0 0.0 1.0  a b eps Mo

32 32      Lx Ly
256 256    Mx My

500        t
0.5        dt
100        t_sample

1          iconfig
0.1 0.05   A_10 A_11
0.2        A_tri

1          iseed
1          ishape
16.0 16.0 9.0    h k roa
"""


def test_read_input_reference_format(tmp_path):
    f = tmp_path / "pfc.input"
    f.write_text(PFC_INPUT)
    p = X.read_input(str(f))
    assert p.n == (256, 256) and p.L == (32.0, 32.0)
    assert p.a == 1.0 and p.b == pytest.approx(1.0 / 3.0)  # read_input.cu:28
    assert p.psi0 == 0.05 and p.sigma == 0.1 and p.dt == 0.5 and p.t_sample == 100
    assert p.iconfig == 1 and p.A_10 == 0.1 and p.A_11 == 0.05 and p.A_tri == 0.2
    assert (p.iseed, p.ishape, p.h, p.k, p.roa) == (1, 1, 16.0, 16.0, 9.0)
    assert X.reference_steps(p) == 1001  # t/dt + 1


def test_read_input_triangular_box_truncates(tmp_path):
    f = tmp_path / "pfc.input"
    f.write_text(PFC_INPUT.replace("1          iconfig", "0          iconfig"))
    p = X.read_input(str(f))
    assert p.L == (float(int(32 * 2 / math.sqrt(3))),) * 2  # initialize.cu:25


def test_read_input_errors(tmp_path):
    with pytest.raises(X.XPFCError) as e:
        X.read_input(str(tmp_path / "missing.input"))
    assert e.value.status == 2
    f = tmp_path / "bad.input"
    f.write_text(PFC_INPUT.replace("1          iconfig", "7          iconfig"))
    with pytest.raises(X.XPFCError) as e:
        X.read_input(str(f))
    assert e.value.status == 3
    f.write_text("1\n0.05\n")
    with pytest.raises(X.XPFCError) as e:
        X.read_input(str(f))
    assert e.value.status == 3


def test_read_input_extensions(tmp_path):
    f = tmp_path / "pfc.input"
    f.write_text(PFC_INPUT + "dim 3\nLz 16\nMz 128\nprecision fp32\ndcf max\npeaks 1 6.283185307179586 1.0 1.0\n")
    p = X.read_input(str(f))
    assert p.n == (256, 256, 128) and p.L[2] == 16.0
    assert p.precision == X.XPFC_FP32_MIXED and p.dcf_mode == X.XPFC_DCF_MAX
    assert len(p.peaks) == 1 and p.peaks[0][2] == 1.0


def test_batch_rejects_mismatched_grid():
    """xpfc_create_batch checks the instances' grids before touching a device."""
    from cases import make
    _, p = make((32, 32))
    _, q = make((64, 32))
    with pytest.raises(X.XPFCError):
        X.BatchContext([p, q])
    _, r = make((32, 32), prec=X.XPFC_FP32_MIXED)
    with pytest.raises(X.XPFCError):
        X.BatchContext([p, r])
    with pytest.raises(ValueError):
        X.BatchContext([])


def _vtk3_text(psi, L):
    """A 3-D legacy VTK file laid out like the 2-D writer (x fastest)."""
    nx, ny, nz = psi.shape
    lines = ["# vtk DataFile Version 3.0", "Density Data", "ASCII", "DATASET STRUCTURED_POINTS",
             "DIMENSIONS %d %d %d" % (nx, ny, nz), "ORIGIN 0 0 0",
             "SPACING %g %g %g" % (L[0] / nx, L[1] / ny, L[2] / nz),
             "POINT_DATA %d" % psi.size, "SCALARS DensityData float", "LOOKUP_TABLE default"]
    lines += ["%g" % v for v in psi.transpose(2, 1, 0).reshape(-1)]
    return "\n".join(lines) + "\n"


def test_read_vtk_2d_and_3d(tmp_path):
    rng = np.random.default_rng(0)
    psi2 = 0.05 + 0.01 * rng.standard_normal((6, 4))
    f2 = tmp_path / "a.vtk"
    f2.write_text(O.vtk_text(psi2, 12, 8))
    dims, sp, v = X.api.read_vtk(str(f2))
    assert dims == (6, 4, 1) and sp == (2.0, 2.0, 1.0)
    # file order is x fastest
    expect = np.array([float("%g" % x) for x in psi2.T.reshape(-1)])
    assert np.array_equal(v, expect)
    psi3 = rng.standard_normal((4, 3, 2))
    f3 = tmp_path / "b.vtk"
    f3.write_text(_vtk3_text(psi3, (4.0, 3.0, 2.0)))
    dims, sp, v = X.api.read_vtk(str(f3))
    assert dims == (4, 3, 2) and sp == (1.0, 1.0, 1.0)
    assert np.array_equal(v.reshape(2, 3, 4).transpose(2, 1, 0),
                          np.vectorize(lambda x: float("%g" % x))(psi3))
    (tmp_path / "c.vtk").write_text("# vtk\nx\n")
    with pytest.raises(X.XPFCError):
        X.api.read_vtk(str(tmp_path / "c.vtk"))


def test_read_input_restart_box_from_vtk(tmp_path, monkeypatch):
    """iconfig 10 takes the box from the restart file (read_input.cu:124-157)."""
    text = PFC_INPUT.replace("1          iconfig", "10         iconfig")
    assert text != PFC_INPUT
    (tmp_path / "pfc.input").write_text(text)
    psi = np.zeros((64, 32))
    (tmp_path / "initial_density_field.vtk").write_text(O.vtk_text(psi, 40, 24))
    monkeypatch.chdir(tmp_path)
    p = X.read_input(str(tmp_path / "pfc.input"))
    assert p.iconfig == 10 and p.L[:2] == (40.0, 24.0)


@pytest.mark.parametrize("n,L,msg", [
    ((56, 32), (7.0, 4.0), "3*2^k"),             # 7 * 8: no compiled line length
    ((36, 32), (4.5, 4.0), "3*2^k"),             # 9 * 4: a factor 3 twice
    ((10, 32), (1.25, 4.0), "3*2^k"),            # 5 * 2: its half line (5) is not compiled
    ((4, 32), (0.5, 4.0), "3*2^k"),              # below the smallest line
    ((4096, 32), (512.0, 4.0), "3*2^k"),         # above the longest line
    ((32, 32), (0.0, 4.0), "box length"),        # empty box
])
def test_create_rejects_bad_grids_before_touching_a_device(n, L, msg):
    p = X.Params(n=n, L=L)
    with pytest.raises(X.XPFCError) as e:
        X.Context(p)
    assert e.value.status == X.api.XPFC_ERR_ARG and msg in str(e.value)


def test_instance_bytes_counts_three_spectra():
    p32 = X.Params(n=(64, 64, 64), L=(8.0, 8.0, 8.0), precision=X.XPFC_FP32_MIXED)
    p64 = X.Params(n=(64, 64, 64), L=(8.0, 8.0, 8.0), precision=X.XPFC_FP64)
    spectrum32 = 64 * 64 * 33 * 8
    b32, b64 = X.instance_bytes(p32), X.instance_bytes(p64)
    assert 3 * spectrum32 <= b32 < 3 * spectrum32 + (1 << 20)
    assert 3 * 2 * spectrum32 <= b64 < 3 * 2 * spectrum32 + (1 << 20)
    assert X.instance_bytes(X.Params(n=(56, 56), L=(7.0, 7.0))) == 0
    # 3*2^k and 5*2^k extents are sized like any other
    spectrum48 = 48 * 25 * 16
    assert 3 * spectrum48 <= X.instance_bytes(X.Params(n=(48, 48), L=(6.0, 6.0))) < 3 * spectrum48 + (1 << 20)
    assert X.instance_bytes(X.Params(n=(2048, 2048), L=(256.0, 256.0))) >= 3 * 2048 * 1025 * 16


def test_read_vtk_rejects_short_or_garbled_data(tmp_path):
    psi = 0.05 + 0.01 * np.random.default_rng(1).standard_normal((6, 4))
    text = O.vtk_text(psi, 12, 8)
    lines = text.splitlines()
    (tmp_path / "short.vtk").write_text("\n".join(lines[:-3]) + "\n")        # 3 values missing
    (tmp_path / "nan.vtk").write_text("\n".join(lines[:-1] + ["abc"]) + "\n")  # a non-number
    (tmp_path / "dims.vtk").write_text(text.replace("DIMENSIONS 6 4 1", "DIMENSIONS 6 -4 1"))
    for name in ("short.vtk", "nan.vtk", "dims.vtk", "missing.vtk"):
        with pytest.raises(X.XPFCError):
            X.api.read_vtk(str(tmp_path / name))


def test_params_layout_and_powf_extension(tmp_path):
    import ctypes
    assert X.lib().xpfc_params_size() == ctypes.sizeof(X.api._Params)
    f = tmp_path / "pfc.input"
    f.write_text(PFC_INPUT + "powf_fp32 1\nint_Mo_residual 0\ndcf max\n")
    p = X.read_input(str(f))
    assert p.powf_fp32 == 1 and p.int_Mo_residual == 0 and p.dcf_mode == X.XPFC_DCF_MAX
    assert X.Params.from_c(p.to_c()).powf_fp32 == 1
    assert X.Params().powf_fp32 == 0
