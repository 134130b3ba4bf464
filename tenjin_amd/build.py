"""Build step of the gfx950 library: compiles tenjin_amd/csrc/*.hip (the
FP64 / mixed forms and the tangent-linear form) into
tenjin_amd/libtenjin.so with hipcc (in-tree, so the built library travels with
the source tree), the C++ host interface tenjin_amd/csrc/host/ into
tenjin_amd/libtenjin_host.so, and the C++ time-loop driver used by the tests
(tests/host/main_loop.cpp) into tests/host/main_loop, the libtorch
autograd Function over the tangent-linear form (csrc/torch/) into
tenjin_amd/tenjin_torch.so, and the device-peak kernels (csrc/tj_peaks.hip)
into tenjin_amd/libtenjin_peaks.so.

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
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtenjin.so")
HOST_LIB = os.path.join(HERE, "libtenjin_host.so")
DRIVER = os.path.join(ROOT, "tests", "host", "main_loop")
INCLUDE = os.path.join(ROOT, "include")
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
    """(source, extra flags, object name): tj_lib.hip twice, the second time
    with the state view as one FAB per box (TJ_FAB_VIEWS)."""
    lib = os.path.join(CSRC, "tj_lib.hip")
    return [(lib, [], "tj_lib.o"), (lib, ["-DTJ_FAB_VIEWS"], "tj_lib_fabs.o"),
            (os.path.join(CSRC, "tj_tangent.hip"), [], "tj_tangent.o")]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose: bool = True, extra=None) -> str:
    """Each translation unit compiles in its own process (objects in a
    temporary folder), then one link makes the shared library."""
    import tempfile

    comp = [f for f in FLAGS if f != "-shared"] + list(extra or [])
    with tempfile.TemporaryDirectory(prefix="tenjin_build_") as tmp:
        objs, procs = [], []
        for src, defs, name in sources():
            obj = os.path.join(tmp, name)
            cmd = [HIPCC] + comp + defs + ["-c", src, "-I", INCLUDE, "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append((cmd, subprocess.Popen(cmd)))
            objs.append(obj)
        failed = [cmd for cmd, p in procs if p.wait() != 0]
        if failed:
            raise subprocess.CalledProcessError(1, failed[0])
        _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", LIB + ".tmp"], verbose)
    os.replace(LIB + ".tmp", LIB)
    return LIB


PEAKS_LIB = os.path.join(HERE, "libtenjin_peaks.so")


def build_peaks(verbose: bool = True) -> str:
    """The device-peak measurement kernels (csrc/tj_peaks.hip ->
    tenjin_amd/libtenjin_peaks.so, loaded by bench.py), with the library's
    flags."""
    src = os.path.join(CSRC, "tj_peaks.hip")
    _run([HIPCC] + FLAGS + [src, "-o", PEAKS_LIB + ".tmp"], verbose)
    os.replace(PEAKS_LIB + ".tmp", PEAKS_LIB)
    return PEAKS_LIB


TORCH_LIB = os.path.join(HERE, "tenjin_torch.so")


def build_torch(verbose: bool = True) -> str:
    """The libtorch autograd Function over the tangent-linear form
    (csrc/torch/tj_autograd.cpp -> tenjin_amd/tenjin_torch.so, a Python
    extension module that links libtenjin.so from its own folder), with the
    flags torch.utils.cpp_extension uses for a C++ extension."""
    import sysconfig

    from torch.utils import cpp_extension as ce

    src = os.path.join(CSRC, "torch", "tj_autograd.cpp")
    cmd = ["c++", "-O2", "-std=c++17", "-fPIC", "-shared", "-DTORCH_EXTENSION_NAME=tenjin_torch",
           "-DTORCH_API_INCLUDE_EXTENSION_H", "-D__HIP_PLATFORM_AMD__=1", "-DUSE_ROCM=1", "-DHIPBLAS_V2"]
    for p in ce.include_paths():
        cmd += ["-isystem", p]
    cmd += ["-isystem", sysconfig.get_paths()["include"], "-I", INCLUDE, src]
    for p in ce.library_paths():
        cmd += ["-L", p]
    cmd += ["-lc10", "-ltorch_cpu", "-ltorch", "-ltorch_python", "-L", HERE, "-ltenjin", "-Wl,-rpath,$ORIGIN",
            "-o", TORCH_LIB + ".tmp"]
    _run(cmd, verbose)
    os.replace(TORCH_LIB + ".tmp", TORCH_LIB)
    return TORCH_LIB


HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result"]


def build_host(verbose: bool = True) -> str:
    """The C++ host interface (host code only; links libtenjin.so)."""
    src = os.path.join(CSRC, "host", "tenjin_host.cpp")
    cmd = [HIPCC] + HOST_FLAGS + ["-shared", "-I", INCLUDE, src, "-L", HERE, "-ltenjin",
                                  "-Wl,-rpath,$ORIGIN", "-o", HOST_LIB + ".tmp"]
    _run(cmd, verbose)
    os.replace(HOST_LIB + ".tmp", HOST_LIB)
    src = os.path.join(ROOT, "tests", "host", "main_loop.cpp")
    cmd = [HIPCC] + HOST_FLAGS + ["-I", INCLUDE, "-I", os.path.join(CSRC, "host"), src, "-L", HERE, "-ltenjin_host",
                                  "-ltenjin", "-Wl,-rpath,$ORIGIN/../../tenjin_amd", "-o", DRIVER + ".tmp"]
    _run(cmd, verbose)
    os.replace(DRIVER + ".tmp", DRIVER)
    return HOST_LIB


if __name__ == "__main__":
    build(extra=sys.argv[1:])
    build_host()
    build_torch()
