"""Builds libmittet_fdtd.so in-tree with hipcc for gfx950.

The library is the only native product artefact: hand-written HIP kernels plus
the C-ABI host layer.  It is compiled with -ffp-contract=off so that every
multiply and add rounds exactly where the reference's C code rounds.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libmittet_fdtd.so")
SOURCES = [os.path.join(HERE, "csrc", f)
           for f in ("fdtd_kernels.hip", "fdtd_adjoint.hip", "fdtd_compat.hip", "fdtd_transform.hip",
                     "fdtd_fortran.hip", "fdtd_api.hip")]
HEADERS = [os.path.join(HERE, "csrc", "fdtd_dev.h"), os.path.join(HERE, "csrc", "fdtd_transform.h"),
           os.path.join(ROOT, "include", "mittet_fdtd.h")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def build(verbose: bool = False) -> str:
    cmd = [HIPCC, *FLAGS, "-o", LIB + ".tmp", *SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libmittet_fdtd.so")
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
