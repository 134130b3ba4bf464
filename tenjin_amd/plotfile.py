"""Reader and writer of one-level AMReX plotfiles ("HyperCLaw-V1.1": Header,
Level_0/Cell_H, Level_0/Cell_D_*), the format the reference writes with
WriteSingleLevelPlotfile (main.cpp:104-107, 173-180, 195-199) and the C++ host
interface writes with amrex_lite::WriteSingleLevelPlotfile.  Host-side I/O only."""
from __future__ import annotations

import os
import re

import numpy as np

_REAL_DESC = "FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))"
_BOX = re.compile(r"\(\(([-\d,]+)\) \(([-\d,]+)\) \(([-\d,]+)\)\)")


def _box(text):
    m = _BOX.search(text)
    lo = [int(v) for v in m.group(1).split(",")]
    hi = [int(v) for v in m.group(2).split(",")]
    return lo, hi


def read_plotfile(path: str):
    """Returns (meta, data): meta has names, dim, time, step, prob_lo/hi, dx,
    domain, boxes; data is [ncomp][nz][ny][nx] float64 over the domain."""
    with open(os.path.join(path, "Header")) as f:
        lines = [ln.rstrip("\n") for ln in f]
    if lines[0] != "HyperCLaw-V1.1":
        raise ValueError(f"{path}: not a HyperCLaw-V1.1 plotfile")
    nc = int(lines[1])
    names = lines[2:2 + nc]
    q = 2 + nc
    dim = int(lines[q])
    time = float(lines[q + 1])
    finest = int(lines[q + 2])
    if finest != 0:
        raise ValueError("single-level plotfiles only")
    prob_lo = [float(v) for v in lines[q + 3].split()]
    prob_hi = [float(v) for v in lines[q + 4].split()]
    dom_lo, dom_hi = _box(lines[q + 6])
    step = int(lines[q + 7])
    dx = [float(v) for v in lines[q + 8].split()]
    meta = dict(names=names, dim=dim, time=time, step=step, prob_lo=prob_lo, prob_hi=prob_hi, dx=dx,
                domain=(dom_lo, dom_hi))
    with open(os.path.join(path, "Level_0", "Cell_H")) as f:
        ch = f.read().split("\n")
    nboxes = int(ch[4].strip("(").split()[0])
    boxes = [_box(ch[5 + b]) for b in range(nboxes)]
    onDisk = [ln.split() for ln in ch if ln.startswith("FabOnDisk:")]
    meta["boxes"] = boxes
    n = [dom_hi[d] - dom_lo[d] + 1 for d in range(dim)] + [1] * (3 - dim)
    data = np.zeros((nc, n[2], n[1], n[0]))
    for (lo, hi), (_, fname, off) in zip(boxes, onDisk):
        with open(os.path.join(path, "Level_0", fname), "rb") as f:
            f.seek(int(off))
            header = f.readline().decode()
            if not header.startswith(_REAL_DESC):
                raise ValueError(f"unsupported FAB encoding: {header[:80]}")
            blo, bhi = _box(header)
            ncomp = int(header.rsplit(")", 1)[1])
            ext = [bhi[d] - blo[d] + 1 for d in range(dim)] + [1] * (3 - dim)
            arr = np.frombuffer(f.read(8 * ncomp * ext[0] * ext[1] * ext[2]), dtype="<f8")
            arr = arr.reshape(ncomp, ext[2], ext[1], ext[0])
            o = [blo[d] - dom_lo[d] for d in range(dim)] + [0] * (3 - dim)
            data[:, o[2]:o[2] + ext[2], o[1]:o[1] + ext[1], o[0]:o[0] + ext[0]] = arr
    return meta, data


def write_plotfile(path: str, data: np.ndarray, names, boxes, L, time: float, step: int):
    """Writes data [ncomp][nz][ny][nx] as a one-level plotfile with the given
    boxes ((lo, hi) index triples over the domain) and domain lengths L."""
    nc = data.shape[0]
    nz, ny, nx = data.shape[1:]
    dim = 1 + (ny > 1 or nz > 1) + (nz > 1)
    n = [nx, ny, nz][:dim]
    dx = [L[d] / n[d] for d in range(dim)]
    tup = lambda v: "(" + ",".join(str(int(x)) for x in v[:dim]) + ")"
    bstr = lambda lo, hi: "(" + tup(lo) + " " + tup(hi) + " " + tup([0, 0, 0]) + ")"
    os.makedirs(os.path.join(path, "Level_0"), exist_ok=True)
    r = lambda v: repr(float(v))
    with open(os.path.join(path, "Header"), "w") as f:
        f.write("HyperCLaw-V1.1\n%d\n" % nc)
        for nm in names[:nc]:
            f.write(nm + "\n")
        f.write(f"{dim}\n{r(time)}\n0\n")
        f.write(" ".join("0.0" for _ in range(dim)) + "\n")
        f.write(" ".join(r(L[d]) for d in range(dim)) + "\n\n")
        f.write(bstr([0, 0, 0], [x - 1 for x in n]) + "\n" + f"{step}\n")
        f.write(" ".join(r(v) for v in dx) + "\n0\n0\n")
        f.write(f"0 {len(boxes)} {r(time)}\n{step}\n")
        for lo, hi in boxes:
            for d in range(dim):
                f.write(f"{r(lo[d] * dx[d])} {r((hi[d] + 1) * dx[d])}\n")
        f.write("Level_0/Cell\n")
    offs, mins, maxs = [], [], []
    with open(os.path.join(path, "Level_0", "Cell_D_00000"), "wb") as f:
        for lo, hi in boxes:
            offs.append(f.tell())
            f.write((_REAL_DESC + bstr(lo, hi) + f" {nc}\n").encode())
            sub = data[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]
            f.write(np.ascontiguousarray(sub, dtype="<f8").tobytes())
            mins.append(sub.reshape(nc, -1).min(axis=1))
            maxs.append(sub.reshape(nc, -1).max(axis=1))
    with open(os.path.join(path, "Level_0", "Cell_H"), "w") as f:
        f.write(f"1\n1\n{nc}\n0\n({len(boxes)} 0\n")
        for lo, hi in boxes:
            f.write(bstr(lo, hi) + "\n")
        f.write(f")\n{len(boxes)}\n")
        for o in offs:
            f.write(f"FabOnDisk: Cell_D_00000 {o}\n")
        for arr in (mins, maxs):
            f.write(f"\n{len(boxes)},{nc}\n")
            for row in arr:
                f.write("".join(r(v) + "," for v in row) + "\n")
