"""Build libxpfc.so (HIP kernels for gfx950 + C++ host runtime) in-tree.

Usage: python -m xpfc_cuda_amd.build   (or xpfc_cuda_amd.build.build())
Output: xpfc_cuda_amd/lib/libxpfc.so and xpfc_cuda_amd/bin/xpfc_run (CLI).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
BINDIR = os.path.join(HERE, "bin")
OBJDIR = os.path.join(HERE, "_build")

ARCH = os.environ.get("XPFC_OFFLOAD_ARCH", "gfx950")
# packed FP32 instructions stay off (they are generated by -O3 otherwise);
# multiply-adds are fused only within one source expression (-ffp-contract=on),
# so the rounding of a line's arithmetic does not depend on the code around it
# (pass variants that restructure a kernel give bitwise-identical fields)
HIPFLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=on",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I", os.path.join(ROOT, "include")]

# the step passes of each line length compile in their own units (xpfc_len_*.hip), in parallel
LEN_UNITS = ["xpfc_len_p2a.hip", "xpfc_len_p2b.hip", "xpfc_len_p2c.hip", "xpfc_len_p2d.hip",
             "xpfc_len_r3a.hip", "xpfc_len_r3b.hip", "xpfc_len_r5a.hip", "xpfc_len_r5b.hip"]
RAG_UNITS = ["xpfc_rag_p2.hip", "xpfc_rag_r3.hip", "xpfc_rag_r5.hip"]
SOURCES = LEN_UNITS + RAG_UNITS + ["xpfc_passes.hip", "xpfc_ops.hip", "xpfc_params.cpp", "xpfc_context.cpp",
                       "xpfc_sched.cpp", "xpfc_service.cpp"]


def _hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l
                    and "is not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build(verbose: bool = False) -> str:
    hipcc = _hipcc()
    os.makedirs(OBJDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(BINDIR, exist_ok=True)
    objs = []
    jobs = []
    for s in SOURCES:
        o = os.path.join(OBJDIR, s + ".o")
        objs.append(o)
        jobs.append([hipcc, *HIPFLAGS, "-c", os.path.join(CSRC, s), "-o", o])
    with ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    lib = os.path.join(LIBDIR, "libxpfc.so")
    tmp = lib + ".tmp"
    _run([hipcc, *HIPFLAGS, "-shared", "-pthread", "-o", tmp, *objs])
    os.replace(tmp, lib)
    cli_src = os.path.join(CSRC, "xpfc_run.cpp")
    if os.path.exists(cli_src):
        cli = os.path.join(BINDIR, "xpfc_run")
        _run([hipcc, *HIPFLAGS, cli_src, "-o", cli, "-L", LIBDIR, "-lxpfc",
              "-Wl,-rpath,$ORIGIN/../lib"])
    return lib


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
