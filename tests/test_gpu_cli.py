"""The command-line driver: reference log lines and VTK files, and a run
resumed from a checkpoint that ends exactly where the straight run ends."""
import os
import subprocess

import numpy as np
import pytest

import xpfc_cuda_amd as X

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(X.library_path()), "..", "bin", "xpfc_run")

PFC = """1          no_val
0.05 0.17   psi_0 s
1.0 3.0 0.0 1.0  a b eps Mo

8 8        Lx Ly
64 64      Mx My

60         t
0.5        dt
50         t_sample

2          iconfig
0.0 0.0    A_10 A_11
0.0        A_tri

0          iseed
0          ishape
0 0 0      h k roa
"""


def _run(cwd, *args):
    r = subprocess.run([CLI, "pfc.input", *args], cwd=cwd, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    return [l for l in r.stdout.splitlines() if l.startswith("iteration:")]


def test_cli_log_lines_and_vtk(tmp_path):
    (tmp_path / "pfc.input").write_text(PFC)
    lines = _run(tmp_path, "--noise", "0.01", "7")
    # t/dt + 1 = 121 steps, FE/RMSE every 50 (main.cu:286,326,412,450)
    p = X.read_input(str(tmp_path / "pfc.input"))
    assert X.reference_steps(p) == 121
    with X.Context(p) as c:
        c.initialize_noise(0.01, 7)
        c.step(121)
        expect = [s.line().rstrip("\n").rstrip() for s in c.samples()]
        F, _ = c.fe_calc()
    assert [l.rstrip() for l in lines[:-1]] == expect
    assert lines[-1] == "iteration: 121 F = %0.10f" % F
    assert lines[0].startswith("iteration: 0 F = ")
    assert "SSE_gpu" in lines[1] and lines[1].startswith("iteration: 50 F = ")
    assert (tmp_path / "density_data0.vtk").exists() and (tmp_path / "density_data121.vtk").exists()


def test_cli_resume_from_checkpoint(tmp_path):
    (tmp_path / "pfc.input").write_text(PFC)
    straight = _run(tmp_path, "--noise", "0.01", "7", "--no-vtk", "--steps", "90")
    _run(tmp_path, "--noise", "0.01", "7", "--no-vtk", "--steps", "40", "--checkpoint", "c.ckpt")
    resumed = _run(tmp_path, "--no-vtk", "--steps", "50", "--resume", "c.ckpt")
    assert resumed[-1] == straight[-1]
    assert resumed[-1].startswith("iteration: 90 F = ")


LATTICE = """1          no_val
0.05 {s}    psi_0 s
1.0 3.0 0.0 1.0  a b eps Mo

{lx} {ly}        Lx Ly
{mx} {my}      Mx My

30         t
0.5        dt
20         t_sample

{iconfig}          iconfig
0.02 0.01  A_10 A_11
0.02       A_tri

0          iseed
0          ishape
0 0 0      h k roa
"""


def _numbers(lines):
    """(step, F, SSE or None) of every printed record."""
    out = []
    for l in lines:
        t = l.split()
        sse = float(t[t.index("SSE_gpu") + 2]) if "SSE_gpu" in t else None
        out.append((int(t[1]), float(t[4]), sse))
    return out


@pytest.mark.parametrize("iconfig,sigma,L,n", [(1, 0.18, (8.0, 8.0), (64, 64)), (0, 0.19, (9.0, 9.0), (64, 64)),
                                             (1, 0.18, (12.0, 10.0), (96, 80))])
def test_cli_lines_match_the_oracle(tmp_path, iconfig, sigma, L, n):
    """xpfc_run on a pfc.input with a lattice initial condition (square, or
    triangular with the box scaled by a_tri and truncated to int: 8 -> 9) and
    the reference's hard-coded DCF (sum of the peaks 2pi and 2sqrt2 pi, rho*B =
    4 and 4/sqrt2): every printed F and SSE_gpu value equals the FP64 oracle's
    run of the same input to the printed digits (F: 10 decimals; SSE: 6
    significant digits).  The third case is a 96 x 80 grid (3*2^k and 5*2^k
    lines) in a 12 x 10 box."""
    from oracle import xpfc_oracle as O
    lx = 8 if L[0] in (8.0, 9.0) else int(L[0])
    ly = 8 if L[1] in (8.0, 9.0) else int(L[1])
    (tmp_path / "pfc.input").write_text(LATTICE.format(s=sigma, iconfig=iconfig, lx=lx, ly=ly, mx=n[0], my=n[1]))
    lines = _run(tmp_path, "--no-vtk")
    op = O.Params(n=n, L=L, sigma=sigma, psi0=0.05, a=1.0, b=1.0 / 3.0, eps=0.0,
                  M=1.0, dt=0.5, peaks=O.reference_peaks(), dcf_mode="sum", t_sample=20)
    psi = O.lattice_ic(op, iconfig, A_10=0.02, A_11=0.01, A_tri=0.02)
    _, rs = O.run(psi, op, 61, sample=True)
    fe = {}
    sse = {}
    for s in rs:
        (fe if s.kind == "FE" else sse)[s.step] = s.value
    got = _numbers(lines)
    # FE + SSE at 0, 20, 40, 60 (one line each), the closing FE at 61
    assert [g[0] for g in got] == [0, 20, 40, 60, 61]
    for step, F, e in got:
        assert abs(F - fe[step]) <= 1e-10 + 1e-11 * abs(fe[step]), (step, F, fe[step])
        if e is not None:
            assert abs(e - sse[step]) <= 1e-5 * abs(sse[step]), (step, e, sse[step])
    assert sum(g[2] is not None for g in got) == 4
