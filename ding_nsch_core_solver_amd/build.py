"""Build the HIP engine (libppsolve.so) in-tree for gfx950.

Usage: ``python -m ding_nsch_core_solver_amd.build`` or ``build()`` from Python.
The library is written next to this file so it travels with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
INCLUDE = os.path.join(REPO_DIR, "include")
LIB_NAME = "libppsolve.so"
LIB_PATH = os.path.join(PKG_DIR, LIB_NAME)

SOURCES = ["pp_fp64.hip", "pp_adjoint.hip", "pp_correct.hip", "pp_bicgstab.hip", "pp_momentum.hip", "pp_ch.hip", "pp_host.cpp"]
ARCH = os.environ.get("PP_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off keeps the reference's unfused double arithmetic (x86-64 baseline, no FMA).
# -packed-fp32-ops keeps the compiler from forming packed FP32 instructions.
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared", "-x", "hip",
    "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise FileNotFoundError("hipcc not found (ROCm toolchain required to build the engine)")


def build(verbose: bool = False) -> str:
    cmd = [hipcc(), *HIPCC_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}"]
    cmd += [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB_PATH + ".tmp"
    cmd += ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"hipcc failed ({proc.returncode})")
    if verbose and proc.stderr.strip():
        sys.stderr.write(proc.stderr)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
