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
