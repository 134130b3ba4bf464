"""Runs the C++ time-loop driver (tests/host/main_loop, built from
tests/host/main_loop.cpp over the C++ host interface) on a problem and reads
its result back."""
from __future__ import annotations

import ctypes
import os
import struct
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "host", "main_loop")


def write_input(path: str, pb, final_time: float, max_steps: int):
    from tenjin_amd import level_desc

    d = level_desc(pb)
    prims = np.ascontiguousarray(pb.prims, dtype=np.float64)
    ncomp = prims.shape[0]
    with open(path, "wb") as f:
        f.write(b"TJIN")
        f.write(struct.pack("<i", 1))
        f.write(bytes(d))
        f.write(struct.pack("<dii", final_time, max_steps, int(pb.shift_pressure)))
        f.write(struct.pack("<qq", ncomp, prims[0].size))
        f.write(prims.tobytes())


def read_output(path: str, pb):
    with open(path, "rb") as f:
        assert f.read(4) == b"TJOU"
        t, n, nan_stop, count = struct.unpack("<diiq", f.read(8 + 4 + 4 + 8))
        data = np.frombuffer(f.read(8 * count), dtype=np.float64)
    return t, n, nan_stop, data.reshape((-1,) + pb.shape)


def run(pb, final_time: float, max_steps: int, workdir: str, timeout: float = 300.0, plotroot: str = "",
        fab_per_box: bool = False, box_checks: bool = False):
    if not os.path.exists(DRIVER):
        raise RuntimeError(f"{DRIVER} is missing: build it with `python tenjin_amd/build.py`")
    src = os.path.join(workdir, "in.bin")
    dst = os.path.join(workdir, "out.bin")
    write_input(src, pb, final_time, max_steps)
    flags = (["--fab-per-box"] if fab_per_box else []) + (["--box-checks"] if box_checks else [])
    r = subprocess.run([DRIVER] + flags + [src, dst] + ([plotroot] if plotroot else []), capture_output=True, text=True,
                       timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"main_loop failed ({r.returncode}): {r.stdout[-2000:]} {r.stderr[-2000:]}")
    out = read_output(dst, pb)
    if box_checks:
        line = [ln for ln in r.stdout.splitlines() if ln.startswith("box checks:")]
        assert line and "state equal" in line[0], r.stdout[-2000:]
    return out
