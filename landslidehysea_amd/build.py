"""Build the HIP library in-tree: landslidehysea_amd/liblhysea_hip.so (gfx950)."""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "liblhysea_hip.so")
SOURCES = ["lhysea_capi.hip", "lhysea_batch.hip", "lhysea_fp64.hip", "lhysea_fp32c.hip", "lhysea_adjoint.hip",
           "lhysea_noalloc.hip", "host/lhysea_output.cpp", "host/lhysea_netcdf.cpp"]
ARCH = os.environ.get("LHYSEA_OFFLOAD_ARCH", "gfx950")

# Packed FP32 instructions stay off in every HIP translation unit.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument", "-Wno-pass-failed"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


HOST_SOURCES = ["host/lhysea_main.cpp", "host/lhysea_problem.cpp"]
EXE = os.path.join(HERE, "L-HySEA.exe")


def build_cli(verbose: bool = False) -> str:
    """Host front end L-HySEA.exe (plain C++), linked against liblhysea_hip.so."""
    cmd = ["g++", "-O2", "-std=c++17", "-Wall"] + [os.path.join(CSRC, s) for s in HOST_SOURCES] + \
          ["-o", EXE + ".tmp", "-L" + HERE, "-l:liblhysea_hip.so", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build of L-HySEA.exe failed")
    os.replace(EXE + ".tmp", EXE)
    return EXE


def build(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIP_FLAGS + [os.path.join(CSRC, s) for s in SOURCES] + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc build of liblhysea_hip.so failed")
    os.replace(tmp, LIB)
    build_cli(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
