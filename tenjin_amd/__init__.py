"""tenjin_amd: the per-cell explicit update of a diffuse-interface
multi-material solver on MI355X (gfx950).

The compute path is the HIP library ``libtenjin.so`` (tenjin_amd/csrc, C ABI in
include/tenjin.h).  This module binds it with ctypes; torch provides device
memory and nothing else.  There is no CPU fallback: if the library is missing
or no GPU is visible, the entry points raise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import configs  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtenjin.so")

TJ_BOX_P2C, TJ_BOX_C2P, TJ_BOX_DERIVE = 1 << 8, 1 << 9, 1 << 10
# y-tile shapes of the batch's stage kernels (columns, rows per lane segment)
BATCH_TILES = ((16, 32), (8, 32), (16, 16))
TJ_PREC_FP64 = 0
TJ_PREC_MIXED = 1
TJ_ST_NAN, TJ_ST_WAVESPEED, TJ_ST_ALPHA_NANS, TJ_ST_ALPHA_SUM, TJ_ST_PLASTIC_DIV = 1, 2, 4, 8, 16
TJ_OP_PLASTIC, TJ_OP_CLEAN_V, TJ_OP_CLEAN_ALPHA, TJ_OP_DT, TJ_OP_NANCHECK, TJ_OP_RADIAL, TJ_OP_REACTIVE = 1, 2, 4, 8, 16, 32, 64
TJ_ST_MIX_BRACKET, TJ_ST_MIX_ITER = 32, 64
TJ_ST_TANGENT_NAN = 128
TJ_TP_GAMMA0, TJ_TP_S, TJ_TP_YIELD, TJ_TP_PREF = 0, 1, 2, 3


class Material(ctypes.Structure):
    _fields_ = [
        ("eos", ctypes.c_int),
        ("solid", ctypes.c_int),
        ("plastic", ctypes.c_int),
        ("mixture", ctypes.c_int),
        ("gamma", ctypes.c_double),
        ("pref", ctypes.c_double),
        ("eref", ctypes.c_double),
        ("CV", ctypes.c_double),
        ("rho0", ctypes.c_double),
        ("c0", ctypes.c_double),
        ("s", ctypes.c_double),
        ("G0", ctypes.c_double),
        ("yield_stress", ctypes.c_double),
        ("gamma_mix", ctypes.c_double),
        ("pref_mix", ctypes.c_double),
        ("eref_mix", ctypes.c_double),
        ("CV_mix", ctypes.c_double),
        ("K0", ctypes.c_double),
        ("eos_alpha", ctypes.c_double),
        ("eos_beta", ctypes.c_double),
        ("e1", ctypes.c_double),
        ("e2", ctypes.c_double),
        ("e3", ctypes.c_double),
        ("e4", ctypes.c_double),
        ("e5", ctypes.c_double),
    ]


class LevelDesc(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int),
        ("n", ctypes.c_int * 3),
        ("L", ctypes.c_double * 3),
        ("max_grid_size", ctypes.c_int),
        ("SOLID", ctypes.c_int),
        ("THINC", ctypes.c_int),
        ("PLASTIC", ctypes.c_int),
        ("MUSCL", ctypes.c_int),
        ("THINCbeta", ctypes.c_double),
        ("CFL", ctypes.c_double),
        ("nmat", ctypes.c_int),
        ("bc_lo", ctypes.c_int * 3),
        ("bc_hi", ctypes.c_int * 3),
        ("mat", Material * 4),
        ("RADIAL", ctypes.c_int),
        ("REACTIVE", ctypes.c_int),
    ]


class MFView(ctypes.Structure):
    _fields_ = [
        ("data", ctypes.c_void_p),
        ("jstride", ctypes.c_longlong),
        ("kstride", ctypes.c_longlong),
        ("nstride", ctypes.c_longlong),
        ("ncomp", ctypes.c_int),
        ("pad_", ctypes.c_int),
        ("fabs", ctypes.c_void_p),
    ]


class Fab(ctypes.Structure):
    """tj_fab: one FAB of a MultiFab of one FAB per box."""
    _fields_ = [
        ("data", ctypes.c_void_p),
        ("lo", ctypes.c_int * 3),
        ("hi", ctypes.c_int * 3),
        ("glo", ctypes.c_int * 3),
        ("pad_", ctypes.c_int),
        ("jstride", ctypes.c_longlong),
        ("kstride", ctypes.c_longlong),
        ("nstride", ctypes.c_longlong),
    ]


# entry points declared in include/tenjin.h
EXPORTS = [
    "tj_create", "tj_destroy", "tj_last_error", "tj_stream", "tj_bind_level", "tj_ncomp",
    "tj_comp_start", "tj_compute_dt", "tj_driver_dt", "tj_set_dt", "tj_get_clock", "tj_rk_stage",
    "tj_post_step", "tj_step", "tj_run", "tj_init_from_prims", "tj_c2p", "tj_derive", "tj_status",
    "tj_reset_history", "tj_synchronize", "tj_prof_enable", "tj_prof_report", "tj_prof_only", "tj_prof_sample", "tj_set_tuning",
    "tj_begin_run", "tj_advance_steps", "tj_lincomb", "tj_graph_capture", "tj_graph_launch",
    "tj_fabset_create", "tj_fabset_view", "tj_fabset_destroy", "tj_level_boxes", "tj_box_op",
    "tj_batch_set_tile", "tj_batch_set_stage2d",
    "tj_batch_create", "tj_batch_destroy", "tj_batch_last_error", "tj_batch_reset_history", "tj_batch_set_seglen",
    "tj_batch_init_from_prims", "tj_batch_begin_run", "tj_batch_advance", "tj_batch_graph_capture",
    "tj_batch_graph_launch", "tj_batch_clocks", "tj_batch_synchronize",
]

_lib = None


def load_library(path: str = LIB_PATH):
    """Loads libtenjin.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python tenjin_amd/build.py`")
    # torch ships its own HIP runtime; loading it first makes the library bind
    # to that same runtime (two HIP runtimes in one process cannot share a device)
    import torch  # noqa: F401
    L = ctypes.CDLL(path)
    vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    mfp = ctypes.POINTER(MFView)
    L.tj_create.argtypes = [ctypes.POINTER(vp), ip]
    L.tj_destroy.argtypes = [vp]
    L.tj_destroy.restype = None
    L.tj_last_error.argtypes = [vp]
    L.tj_last_error.restype = ctypes.c_char_p
    L.tj_stream.argtypes = [vp]
    L.tj_stream.restype = vp
    L.tj_bind_level.argtypes = [vp, ctypes.POINTER(LevelDesc)]
    L.tj_ncomp.argtypes = [vp]
    L.tj_comp_start.argtypes = [vp, ip]
    L.tj_compute_dt.argtypes = [vp, mfp, ctypes.POINTER(dp)]
    L.tj_driver_dt.argtypes = [vp, dp]
    L.tj_set_dt.argtypes = [vp, dp, ip]
    L.tj_get_clock.argtypes = [vp, ctypes.POINTER(dp), ctypes.POINTER(dp), ctypes.POINTER(ip)]
    L.tj_rk_stage.argtypes = [vp, ip, mfp, mfp, mfp, ip]
    L.tj_post_step.argtypes = [vp, mfp, ctypes.c_uint]
    L.tj_step.argtypes = [vp, mfp, mfp, ip]
    L.tj_run.argtypes = [vp, mfp, mfp, dp, ip, ip, ip, ctypes.POINTER(dp), ctypes.POINTER(ip)]
    L.tj_init_from_prims.argtypes = [vp, mfp, ip]
    L.tj_c2p.argtypes = [vp, mfp]
    L.tj_derive.argtypes = [vp, mfp]
    L.tj_status.argtypes = [vp, ctypes.POINTER(ctypes.c_uint32), ip]
    L.tj_reset_history.argtypes = [vp]
    L.tj_synchronize.argtypes = [vp]
    L.tj_prof_enable.argtypes = [vp, ip]
    L.tj_prof_report.argtypes = [vp, ctypes.c_char_p, ip]
    L.tj_prof_only.argtypes = [vp, ctypes.c_char_p]
    L.tj_prof_sample.argtypes = [vp, ctypes.c_int]
    L.tj_set_tuning.argtypes = [vp, ip, ip]
    L.tj_begin_run.argtypes = [vp, mfp, dp]
    L.tj_advance_steps.argtypes = [vp, mfp, mfp, ip, ip]
    L.tj_lincomb.argtypes = [vp, mfp, dp, mfp, dp, mfp]
    L.tj_graph_capture.argtypes = [vp, mfp, mfp, ip, ip]
    L.tj_graph_launch.argtypes = [vp]
    L.tj_fabset_create.argtypes = [vp, ip, ctypes.POINTER(Fab), ip, ctypes.POINTER(vp)]
    L.tj_fabset_view.argtypes = [vp, mfp]
    L.tj_fabset_destroy.argtypes = [vp]
    L.tj_fabset_destroy.restype = None
    L.tj_level_boxes.argtypes = [vp, ctypes.POINTER(ip), ctypes.POINTER(ip)]
    L.tj_box_op.argtypes = [vp, mfp, ctypes.POINTER(ip), ctypes.POINTER(ip), ctypes.c_uint, ctypes.POINTER(dp)]
    L.tj_batch_create.argtypes = [ctypes.POINTER(vp), ip, ip, ctypes.POINTER(LevelDesc), mfp, mfp]
    L.tj_batch_destroy.argtypes = [vp]
    L.tj_batch_destroy.restype = None
    L.tj_batch_last_error.argtypes = [vp]
    L.tj_batch_last_error.restype = ctypes.c_char_p
    L.tj_batch_reset_history.argtypes = [vp]
    L.tj_batch_set_seglen.argtypes = [vp, ip, ip]
    L.tj_batch_set_tile.argtypes = [vp, ip, ip]
    L.tj_batch_set_stage2d.argtypes = [vp, ip]
    L.tj_batch_init_from_prims.argtypes = [vp, ip]
    L.tj_batch_begin_run.argtypes = [vp, dp]
    L.tj_batch_advance.argtypes = [vp, ip, ip]
    L.tj_batch_graph_capture.argtypes = [vp, ip, ip]
    L.tj_batch_graph_launch.argtypes = [vp]
    L.tj_batch_clocks.argtypes = [vp, ctypes.POINTER(dp), ctypes.POINTER(ip), ctypes.POINTER(ctypes.c_uint32)]
    L.tj_batch_synchronize.argtypes = [vp]
    L.tj_tangent_create.argtypes = [vp, ip, ctypes.POINTER(ip), ctypes.POINTER(ip), ip, ctypes.POINTER(vp)]
    L.tj_tangent_destroy.argtypes = [vp]
    L.tj_tangent_destroy.restype = None
    L.tj_tangent_run.argtypes = [vp, mfp, ip, mfp, mfp, dp, ip, ip, ip, ip, ip, ctypes.POINTER(dp), ctypes.POINTER(ip)]
    L.tj_tangent_state.argtypes = [vp, ip, mfp]
    _lib = L
    return L


def level_desc(pb) -> LevelDesc:
    d = LevelDesc()
    d.dim = pb.dim
    for a in range(3):
        d.n[a] = pb.n[a] if a < pb.dim else 1
        d.L[a] = pb.L[a]
        d.bc_lo[a] = pb.bc_lo[a]
        d.bc_hi[a] = pb.bc_hi[a]
    d.max_grid_size = pb.max_grid_size
    d.SOLID, d.THINC, d.PLASTIC, d.MUSCL = pb.SOLID, pb.THINC, pb.PLASTIC, pb.MUSCL
    d.RADIAL = pb.RADIAL
    d.REACTIVE = pb.REACTIVE
    d.THINCbeta = pb.THINCbeta
    d.CFL = pb.CFL
    d.nmat = pb.nmat
    for m, M in enumerate(pb.materials):
        t = d.mat[m]
        t.eos, t.solid, t.plastic = M.eos, M.solid, M.plastic
        t.gamma, t.pref, t.eref, t.CV = M.gamma, M.pref, M.eref, M.CV
        t.rho0, t.c0, t.s, t.G0, t.yield_stress = M.rho0, M.c0, M.s, M.G0, M.yield_stress
        t.mixture = M.mixture
        t.gamma_mix, t.pref_mix, t.eref_mix, t.CV_mix = M.gamma2, M.pref2, M.eref2, M.CV2
        t.K0, t.eos_alpha, t.eos_beta = M.K0, M.eos_alpha, M.eos_beta
        t.e1, t.e2, t.e3, t.e4, t.e5 = M.e1, M.e2, M.e3, M.e4, M.e5
    return d


class TenjinError(RuntimeError):
    pass


class GpuLevel:
    """A bound level on one GPU with two state buffers (S: the state, T: the
    stage-1 buffer), each a [ncomp][nz][ny][nx] float64 device array."""

    def __init__(self, pb, device: int = 0, seglen: int | None = None):
        import torch

        if not torch.cuda.is_available():
            raise TenjinError("no GPU visible: the tenjin_amd update runs on the GPU only")
        self.torch = torch
        self.lib = load_library()
        self.pb = pb
        self.ctx = ctypes.c_void_p()
        self._chk(self.lib.tj_create(ctypes.byref(self.ctx), device), "tj_create")
        self._desc = level_desc(pb)
        self._chk(self.lib.tj_bind_level(self.ctx, ctypes.byref(self._desc)), "tj_bind_level")
        self.ncomp = self.lib.tj_ncomp(self.ctx)
        self.shape = pb.shape
        dev = torch.device("cuda", device)
        self.S = torch.zeros((self.ncomp,) + self.shape, dtype=torch.float64, device=dev)
        self.T = torch.zeros_like(self.S)
        self.vS = self._view(self.S)
        self.vT = self._view(self.T)

    def _chk(self, rc, what):
        if rc != 0:
            msg = self.lib.tj_last_error(self.ctx).decode() if self.ctx else ""
            raise TenjinError(f"{what} failed ({rc}): {msg}")

    def _view(self, t) -> MFView:
        nz, ny, nx = self.shape
        v = MFView()
        v.data = t.data_ptr()
        v.jstride = nx
        v.kstride = nx * ny
        v.nstride = nx * ny * nz
        v.ncomp = self.ncomp
        return v

    def close(self):
        if self.ctx:
            self.lib.tj_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # state transfer -------------------------------------------------------
    def set_state(self, arr: np.ndarray):
        self.torch.cuda.synchronize()
        self.S.copy_(self.torch.from_numpy(np.ascontiguousarray(arr, dtype=np.float64)))
        self.torch.cuda.synchronize()

    def get_state(self) -> np.ndarray:
        self.sync()
        return self.S.cpu().numpy()

    def sync(self):
        self._chk(self.lib.tj_synchronize(self.ctx), "tj_synchronize")
        self.torch.cuda.synchronize()

    # entry points ----------------------------------------------------------
    def init_from_prims(self, prims: np.ndarray, shift_pressure: int = 0):
        self.set_state(prims)
        self._chk(self.lib.tj_init_from_prims(self.ctx, ctypes.byref(self.vS), shift_pressure), "tj_init_from_prims")
        self._chk(self.lib.tj_reset_history(self.ctx), "tj_reset_history")
        self.sync()

    def init_from_device(self, prims, shift_pressure: int = 0):
        """init_from_prims from a device tensor shaped like S (a device-to-device
        copy; no host transfer and no synchronisation after the launch)."""
        self.torch.cuda.synchronize()
        self.S.copy_(prims)
        self.torch.cuda.synchronize()
        self._chk(self.lib.tj_init_from_prims(self.ctx, ctypes.byref(self.vS), shift_pressure), "tj_init_from_prims")
        self._chk(self.lib.tj_reset_history(self.ctx), "tj_reset_history")

    def compute_dt(self) -> float:
        dt = ctypes.c_double()
        self._chk(self.lib.tj_compute_dt(self.ctx, ctypes.byref(self.vS), ctypes.byref(dt)), "tj_compute_dt")
        return dt.value

    def set_dt(self, dt: float, reset_clock: int = 0):
        self._chk(self.lib.tj_set_dt(self.ctx, dt, reset_clock), "tj_set_dt")

    def step(self, prec: int = TJ_PREC_FP64):
        self._chk(self.lib.tj_step(self.ctx, ctypes.byref(self.vS), ctypes.byref(self.vT), prec), "tj_step")

    def stage(self, stage: int, prec: int = TJ_PREC_FP64):
        if stage == 1:
            rc = self.lib.tj_rk_stage(self.ctx, 1, ctypes.byref(self.vS), ctypes.byref(self.vS), ctypes.byref(self.vT), prec)
        else:
            rc = self.lib.tj_rk_stage(self.ctx, 2, ctypes.byref(self.vT), ctypes.byref(self.vS), ctypes.byref(self.vS), prec)
        self._chk(rc, "tj_rk_stage")

    def post_step(self, ops: int):
        self._chk(self.lib.tj_post_step(self.ctx, ctypes.byref(self.vS), ops), "tj_post_step")

    def run(self, final_time: float, max_steps: int = 1 << 30, prec: int = TJ_PREC_FP64, check_every: int = 16):
        t = ctypes.c_double()
        n = ctypes.c_int()
        self._chk(self.lib.tj_run(self.ctx, ctypes.byref(self.vS), ctypes.byref(self.vT), final_time, max_steps, prec,
                                  check_every, ctypes.byref(t), ctypes.byref(n)), "tj_run")
        return n.value, t.value

    def begin_run(self, final_time: float = float("inf")):
        self._chk(self.lib.tj_begin_run(self.ctx, ctypes.byref(self.vS), final_time), "tj_begin_run")

    def advance(self, nsteps: int, prec: int = TJ_PREC_FP64):
        self._chk(self.lib.tj_advance_steps(self.ctx, ctypes.byref(self.vS), ctypes.byref(self.vT), nsteps, prec),
                  "tj_advance_steps")

    def graph_capture(self, nsteps: int, prec: int = TJ_PREC_FP64):
        """Records nsteps loop bodies (dt, clamps, step, NaN stop) as one hipGraph."""
        self._chk(self.lib.tj_graph_capture(self.ctx, ctypes.byref(self.vS), ctypes.byref(self.vT), nsteps, prec),
                  "tj_graph_capture")

    def graph_launch(self):
        self._chk(self.lib.tj_graph_launch(self.ctx), "tj_graph_launch")

    def clock(self):
        t, dt, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        self._chk(self.lib.tj_get_clock(self.ctx, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(n)), "tj_get_clock")
        return t.value, dt.value, n.value

    def status(self, clear: int = 0) -> int:
        s = ctypes.c_uint32()
        self._chk(self.lib.tj_status(self.ctx, ctypes.byref(s), clear), "tj_status")
        return s.value

    def set_tuning(self, seglen: int = -1, fuse: int = -1):
        """Rows per marching segment (-1: unchanged) and the form of a stage of a
        two-dimensional fluid level: 2 the stage kernels (default), 1 flux
        planes with the update inside the last directional pass, 0 flux planes
        and a separate update kernel (-1 unchanged)."""
        self._chk(self.lib.tj_set_tuning(self.ctx, seglen, fuse), "tj_set_tuning")

    def prof_enable(self, on: bool = True):
        self._chk(self.lib.tj_prof_enable(self.ctx, int(on)), "tj_prof_enable")

    def prof_only(self, name=None):
        """Time only the launches named `name` (None: every launch)."""
        self._chk(self.lib.tj_prof_only(self.ctx, name.encode() if name else None), "tj_prof_only")

    def prof_sample(self, every: int = 1):
        """Time the first and then every `every`-th qualifying launch."""
        self._chk(self.lib.tj_prof_sample(self.ctx, int(every)), "tj_prof_sample")

    def prof_report(self) -> dict:
        buf = ctypes.create_string_buffer(1 << 16)
        self._chk(self.lib.tj_prof_report(self.ctx, buf, len(buf)), "tj_prof_report")
        out = {}
        for line in buf.value.decode().splitlines():
            name, cnt, ms = line.split()
            out[name] = (int(cnt), float(ms))
        return out

    def derive(self):
        self._chk(self.lib.tj_derive(self.ctx, ctypes.byref(self.vS)), "tj_derive")

    def c2p(self):
        """conservativeToPrimitive of S (tj_c2p)."""
        self._chk(self.lib.tj_c2p(self.ctx, ctypes.byref(self.vS)), "tj_c2p")

    def box_op(self, lo, hi, ops: int) -> float:
        """Cell-wise ops (TJ_BOX_*, TJ_OP_* bits) on the box [lo, hi] of S
        only (tj_box_op); returns the box's dt with TJ_OP_DT."""
        L3 = ctypes.c_int * 3
        dt = ctypes.c_double()
        self._chk(self.lib.tj_box_op(self.ctx, ctypes.byref(self.vS), L3(*lo), L3(*hi), ops, ctypes.byref(dt)),
                  "tj_box_op")
        return dt.value


class GpuFabLevel(GpuLevel):
    """GpuLevel whose S and T are MultiFabs of one FAB per box, as an AMReX
    MultiFab holds them: each box of the level's maxSize decomposition owns a
    float64 device tensor [ncomp][nz + 2g][ny + 2g][nx + 2g + pad] (g ghost
    cells in the active dimensions, `pad` extra doubles per row), and every
    entry point runs on them through tj_fabset views, with no copy into a
    level array.  Ghost cells are never read or written by the library; they
    start as `ghost_fill`.  fabS[b] / fabT[b] and boxes[b] = (lo, hi) are the
    FABs and their boxes."""

    def __init__(self, pb, device: int = 0, ghost: int = 2, pad: int = 0, ghost_fill: float = float("nan")):
        import torch

        if not torch.cuda.is_available():
            raise TenjinError("no GPU visible: the tenjin_amd update runs on the GPU only")
        self.torch = torch
        self.lib = load_library()
        self.pb = pb
        self.ctx = ctypes.c_void_p()
        self._sets = []
        self._chk(self.lib.tj_create(ctypes.byref(self.ctx), device), "tj_create")
        self._desc = level_desc(pb)
        self._chk(self.lib.tj_bind_level(self.ctx, ctypes.byref(self._desc)), "tj_bind_level")
        self.ncomp = self.lib.tj_ncomp(self.ctx)
        self.shape = pb.shape
        self.ghost = ghost
        nb = ctypes.c_int()
        self._chk(self.lib.tj_level_boxes(self.ctx, ctypes.byref(nb), None), "tj_level_boxes")
        lohi = (ctypes.c_int * (6 * nb.value))()
        self._chk(self.lib.tj_level_boxes(self.ctx, ctypes.byref(nb), lohi), "tj_level_boxes")
        self.boxes = [(tuple(lohi[6 * q:6 * q + 3]), tuple(lohi[6 * q + 3:6 * q + 6])) for q in range(nb.value)]
        dev = torch.device("cuda", device)
        self.fabS, self.fabT = [], []
        for fabs in (self.fabS, self.fabT):
            recs = (Fab * len(self.boxes))()
            for q, (lo, hi) in enumerate(self.boxes):
                g = [ghost if d < pb.dim else 0 for d in range(3)]
                ext = [hi[d] - lo[d] + 1 + 2 * g[d] for d in range(3)]
                t = torch.full((self.ncomp, ext[2], ext[1], ext[0] + pad), ghost_fill, dtype=torch.float64, device=dev)
                fabs.append(t)
                r = recs[q]
                r.data = t.data_ptr()
                for d in range(3):
                    r.lo[d], r.hi[d], r.glo[d] = lo[d], hi[d], lo[d] - g[d]
                r.jstride = ext[0] + pad
                r.kstride = r.jstride * ext[1]
                r.nstride = r.kstride * ext[2]
            fs = ctypes.c_void_p()
            self._chk(self.lib.tj_fabset_create(self.ctx, len(self.boxes), recs, self.ncomp, ctypes.byref(fs)),
                      "tj_fabset_create")
            self._sets.append(fs)
        self.vS, self.vT = MFView(), MFView()
        self._chk(self.lib.tj_fabset_view(self._sets[0], ctypes.byref(self.vS)), "tj_fabset_view")
        self._chk(self.lib.tj_fabset_view(self._sets[1], ctypes.byref(self.vT)), "tj_fabset_view")

    def _valid(self, q):
        lo, hi = self.boxes[q]
        g = [self.ghost if d < self.pb.dim else 0 for d in range(3)]
        return (slice(None), slice(g[2], g[2] + hi[2] - lo[2] + 1), slice(g[1], g[1] + hi[1] - lo[1] + 1),
                slice(g[0], g[0] + hi[0] - lo[0] + 1))

    def _level(self, q):
        lo, hi = self.boxes[q]
        return (slice(None), slice(lo[2], hi[2] + 1), slice(lo[1], hi[1] + 1), slice(lo[0], hi[0] + 1))

    def set_state(self, arr: np.ndarray):
        self.torch.cuda.synchronize()
        src = self.torch.from_numpy(np.ascontiguousarray(arr, dtype=np.float64)).to(self.fabS[0].device)
        for q in range(len(self.boxes)):
            self.fabS[q][self._valid(q)] = src[self._level(q)]
        self.torch.cuda.synchronize()

    def get_state(self, which: str = "S") -> np.ndarray:
        self.sync()
        fabs = self.fabS if which == "S" else self.fabT
        out = np.empty((self.ncomp,) + self.shape)
        for q in range(len(self.boxes)):
            out[self._level(q)] = fabs[q][self._valid(q)].cpu().numpy()
        return out

    def ghosts(self):
        """Every FAB's ghost cells (and row padding), flattened: what the
        library must leave as it found it."""
        self.sync()
        parts = []
        for fabs in (self.fabS, self.fabT):
            for q, t in enumerate(fabs):
                m = self.torch.ones_like(t, dtype=self.torch.bool)
                m[self._valid(q)] = False
                parts.append(t[m].cpu().numpy())
        return np.concatenate(parts)

    def init_from_device(self, prims, shift_pressure: int = 0):
        raise TenjinError("GpuFabLevel: use init_from_prims")

    def close(self):
        if self.ctx:
            self.lib.tj_synchronize(self.ctx)
            for fs in self._sets:
                self.lib.tj_fabset_destroy(fs)
            self._sets = []
        super().close()


class GpuBatch:
    """An ensemble of independent fluid problems (1D and 2D) advanced together
    on one GPU through the tj_batch_* entry points; S[p], T[p] are torch
    float64 device tensors [ncomp][nz][ny][nx] per problem."""

    def __init__(self, problems, device: int = 0):
        import torch

        if not torch.cuda.is_available():
            raise TenjinError("no GPU visible: the tenjin_amd update runs on the GPU only")
        self.torch = torch
        self.lib = load_library()
        self.problems = list(problems)
        n = len(self.problems)
        dev = torch.device("cuda", device)
        self.S, self.T = [], []
        self._descs = (LevelDesc * n)()
        vS, vT = (MFView * n)(), (MFView * n)()
        for p, pb in enumerate(self.problems):
            self._descs[p] = level_desc(pb)
            nc = pb.layout.ncomp
            S = torch.zeros((nc,) + pb.shape, dtype=torch.float64, device=dev)
            T = torch.zeros_like(S)
            self.S.append(S)
            self.T.append(T)
            for v, t in ((vS, S), (vT, T)):
                nz, ny, nx = pb.shape
                v[p].data = t.data_ptr()
                v[p].jstride = nx
                v[p].kstride = nx * ny
                v[p].nstride = nx * ny * nz
                v[p].ncomp = nc
        self._vS, self._vT = vS, vT
        self.b = ctypes.c_void_p()
        rc = self.lib.tj_batch_create(ctypes.byref(self.b), device, n, self._descs, vS, vT)
        if rc != 0:
            msg = self.lib.tj_batch_last_error(self.b).decode() if self.b else ""
            self.close()
            raise TenjinError(f"tj_batch_create failed ({rc}): {msg}")

    def _chk(self, rc, what):
        if rc != 0:
            msg = self.lib.tj_batch_last_error(self.b).decode() if self.b else ""
            raise TenjinError(f"{what} failed ({rc}): {msg}")

    def close(self):
        if self.b:
            self.lib.tj_batch_destroy(self.b)
            self.b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_from_prims(self, shift_pressure: int = 0):
        self.torch.cuda.synchronize()
        for S, pb in zip(self.S, self.problems):
            S.copy_(self.torch.from_numpy(np.ascontiguousarray(pb.prims, dtype=np.float64)))
        self.torch.cuda.synchronize()
        self._chk(self.lib.tj_batch_init_from_prims(self.b, shift_pressure), "tj_batch_init_from_prims")
        self._chk(self.lib.tj_batch_reset_history(self.b), "tj_batch_reset_history")

    def set_seglen(self, prob: int, seglen: int):
        self._chk(self.lib.tj_batch_set_seglen(self.b, prob, seglen), "tj_batch_set_seglen")

    def set_tile(self, prob: int, tile: int):
        """y-tile shape of problem prob's stage kernels (BATCH_TILES index)."""
        self._chk(self.lib.tj_batch_set_tile(self.b, prob, tile), "tj_batch_set_tile")

    def set_stage2d(self, on: bool):
        """2D problems on the stage kernels (True, default) or the flux-plane passes."""
        self._chk(self.lib.tj_batch_set_stage2d(self.b, int(on)), "tj_batch_set_stage2d")

    def begin_run(self, final_time: float = float("inf")):
        self._chk(self.lib.tj_batch_begin_run(self.b, final_time), "tj_batch_begin_run")

    def advance(self, nsteps: int, prec: int = TJ_PREC_FP64):
        self._chk(self.lib.tj_batch_advance(self.b, nsteps, prec), "tj_batch_advance")

    def graph_capture(self, nsteps: int, prec: int = TJ_PREC_FP64):
        self._chk(self.lib.tj_batch_graph_capture(self.b, nsteps, prec), "tj_batch_graph_capture")

    def graph_launch(self):
        self._chk(self.lib.tj_batch_graph_launch(self.b), "tj_batch_graph_launch")

    def sync(self):
        self._chk(self.lib.tj_batch_synchronize(self.b), "tj_batch_synchronize")
        self.torch.cuda.synchronize()

    def clocks(self):
        n = len(self.problems)
        t = (ctypes.c_double * n)()
        k = (ctypes.c_int * n)()
        st = (ctypes.c_uint32 * n)()
        self._chk(self.lib.tj_batch_clocks(self.b, t, k, st), "tj_batch_clocks")
        return list(t), list(k), list(st)

    def state(self, p: int) -> np.ndarray:
        self.sync()
        return self.S[p].cpu().numpy()
