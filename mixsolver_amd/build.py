"""Build libmixsolver_hip.so in-tree for gfx950 with hipcc.

    python -m mixsolver_amd.build        (or __graft_entry__.build())
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmixsolver_hip.so")
SOURCES = ["ms_setup.cpp", "ms_linalg.hip", "ms_physics.hip", "ms_serve.hip", "ms_output.cpp", "ms_noalloc.hip", "ms_schur.hip",
           "ms_sweep.cpp", "ms_capi.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
         # no FMA contraction: assembly and preconditioner rows stay bit-exact
         # with the reference's evaluation order
         "-ffp-contract=off",
         # never emit packed FP32 instructions
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]


def build(verbose: bool = False) -> str:
    objs = [os.path.join(SRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    if verbose:
        sys.stderr.write(r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
