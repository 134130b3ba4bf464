"""Build step of the gfx950 library: compiles tenjin_amd/csrc/*.hip into
tenjin_amd/libtenjin.so with hipcc (in-tree, so the built library travels with
the source tree).

Flags:
  --offload-arch=gfx950            MI355X (CDNA4) only
  -ffp-contract=off                no FMA formation: the FP64 form keeps the
                                   rounding of the reference's CPU expressions
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                   no packed FP32 instructions
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtenjin.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def sources():
    return [os.path.join(CSRC, "tj_lib.hip")]


def build(verbose: bool = True, extra=None) -> str:
    cmd = [HIPCC] + FLAGS + list(extra or []) + sources() + ["-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(extra=sys.argv[1:])
