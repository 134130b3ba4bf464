"""Test infrastructure only: ctypes wrapper over the oracle (CPU restatement).

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


class Params(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int),
        ("n", ctypes.c_int * 3),
        ("L", ctypes.c_double * 3),
        ("max_grid_size", ctypes.c_int),
        ("nghost", ctypes.c_int),
        ("SOLID", ctypes.c_int),
        ("THINC", ctypes.c_int),
        ("PLASTIC", ctypes.c_int),
        ("MUSCL", ctypes.c_int),
        ("THINCbeta", ctypes.c_double),
        ("CFL", ctypes.c_double),
        ("nmat", ctypes.c_int),
        ("bc_lo", ctypes.c_int * 3),
        ("bc_hi", ctypes.c_int * 3),
        ("math_mode", ctypes.c_int),
        ("eos", ctypes.c_int * 4),
        ("solid", ctypes.c_int * 4),
        ("plastic", ctypes.c_int * 4),
        ("gamma", ctypes.c_double * 4),
        ("pref", ctypes.c_double * 4),
        ("eref", ctypes.c_double * 4),
        ("CV", ctypes.c_double * 4),
        ("rho0", ctypes.c_double * 4),
        ("c0", ctypes.c_double * 4),
        ("s", ctypes.c_double * 4),
        ("G0", ctypes.c_double * 4),
        ("yield_", ctypes.c_double * 4),
        ("RADIAL", ctypes.c_int),
        ("REACTIVE", ctypes.c_int),
        ("mixture", ctypes.c_int * 4),
        ("gamma2", ctypes.c_double * 4),
        ("pref2", ctypes.c_double * 4),
        ("eref2", ctypes.c_double * 4),
        ("CV2", ctypes.c_double * 4),
        ("K0", ctypes.c_double * 4),
        ("ealpha", ctypes.c_double * 4),
        ("ebeta", ctypes.c_double * 4),
        ("e1", ctypes.c_double * 4),
        ("e2", ctypes.c_double * 4),
        ("e3", ctypes.c_double * 4),
        ("e4", ctypes.c_double * 4),
        ("e5", ctypes.c_double * 4),
    ]


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB):
        subprocess.check_call(["make", "-B" if force else "-s", "-C", HERE])
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
        L.tjo_create.restype = vp
        L.tjo_create.argtypes = [ctypes.POINTER(Params)]
        L.tjo_destroy.argtypes = [vp]
        for f in ("tjo_ncomp", "tjo_nboxes", "tjo_status", "tjo_cleanup_alpha"):
            getattr(L, f).argtypes = [vp]
            getattr(L, f).restype = ctypes.c_int
        L.tjo_comp_start.argtypes = [vp, ctypes.c_int]
        L.tjo_box.argtypes = [vp, ctypes.c_int, ip, ip]
        L.tjo_get_valid.argtypes = [vp, ctypes.c_int, dp]
        L.tjo_set_valid.argtypes = [vp, ctypes.c_int, dp]
        L.tjo_get_fab.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
        L.tjo_init_from_prims.argtypes = [vp, ctypes.c_int]
        L.tjo_timestep.argtypes = [vp]
        L.tjo_timestep.restype = ctypes.c_double
        L.tjo_step.argtypes = [vp, ctypes.c_double]
        L.tjo_geometric_source.argtypes = [vp, ctypes.c_double]
        L.tjo_reactive.argtypes = [vp, ctypes.c_double]
        L.tjo_run.argtypes = [vp, ctypes.c_double, ctypes.c_int, dp]
        L.tjo_run.restype = ctypes.c_int
        for f in ("tjo_cleanup_v", "tjo_plastic", "tjo_c2p", "tjo_p2c", "tjo_sound"):
            getattr(L, f).argtypes = [vp]
        L.tjo_reconstruct.argtypes = [vp, ctypes.c_int]
        L.tjo_face_flux.argtypes = [vp, ctypes.c_int, dp, dp, dp]
        L.tjo_vanleer.argtypes = [ctypes.c_double]
        L.tjo_vanleer.restype = ctypes.c_double
        L.tjo_math.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.tjo_math.restype = ctypes.c_double
        L.tjo_set_threads.argtypes = [ctypes.c_int]
        L.tjo_transverse.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.tjo_transverse.restype = ctypes.c_double
        L.tjo_sqrtm.argtypes = [dp, dp]
        L.tjo_inv3.argtypes = [dp, dp]
        L.tjo_det3.argtypes = [dp]
        L.tjo_det3.restype = ctypes.c_double
        L.tjo_chop.argtypes = [ctypes.c_int, ctypes.c_int, ip]
        L.tjo_chop.restype = ctypes.c_int
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def make_params(pb, math_mode: int = 0) -> Params:
    q = Params()
    q.dim = pb.dim
    for a in range(3):
        q.n[a] = pb.n[a] if a < pb.dim else 1
        q.L[a] = pb.L[a]
        q.bc_lo[a] = pb.bc_lo[a]
        q.bc_hi[a] = pb.bc_hi[a]
    q.max_grid_size = pb.max_grid_size
    q.nghost = pb.nghost
    q.SOLID, q.THINC, q.PLASTIC, q.MUSCL = pb.SOLID, pb.THINC, pb.PLASTIC, pb.MUSCL
    q.RADIAL = pb.RADIAL
    q.REACTIVE = pb.REACTIVE
    q.THINCbeta = pb.THINCbeta
    q.CFL = pb.CFL
    q.nmat = pb.nmat
    q.math_mode = math_mode
    for m, M in enumerate(pb.materials):
        q.eos[m], q.solid[m], q.plastic[m] = M.eos, M.solid, M.plastic
        q.gamma[m], q.pref[m], q.eref[m], q.CV[m] = M.gamma, M.pref, M.eref, M.CV
        q.rho0[m], q.c0[m], q.s[m], q.G0[m], q.yield_[m] = M.rho0, M.c0, M.s, M.G0, M.yield_stress
        q.mixture[m] = M.mixture
        q.gamma2[m], q.pref2[m], q.eref2[m], q.CV2[m] = M.gamma2, M.pref2, M.eref2, M.CV2
        q.K0[m], q.ealpha[m], q.ebeta[m] = M.K0, M.eos_alpha, M.eos_beta
        q.e1[m], q.e2[m], q.e3[m], q.e4[m], q.e5[m] = M.e1, M.e2, M.e3, M.e4, M.e5
    return q


class Oracle:
    """One restated level (the nine CellArrays of main.cpp:64-72)."""

    U1, U, U2, UL, UR, ULSTAR, URSTAR, USTARSTAR, GRAD = range(9)

    def __init__(self, pb, math_mode: int = 0, init: bool = True):
        self.pb = pb
        self._params = make_params(pb, math_mode)
        self.h = lib().tjo_create(ctypes.byref(self._params))
        self.ncomp = lib().tjo_ncomp(self.h)
        self.shape = pb.shape
        if init and pb.prims is not None:
            self.set(pb.prims)
            lib().tjo_init_from_prims(self.h, pb.shift_pressure)

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib().tjo_destroy(h)
            self.h = None

    def get(self, which: int = 0) -> np.ndarray:
        out = np.empty((self.ncomp,) + self.shape, dtype=np.float64)
        lib().tjo_get_valid(self.h, which, _dp(out))
        return out

    def set(self, arr: np.ndarray, which: int = 0):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == (self.ncomp,) + self.shape
        lib().tjo_set_valid(self.h, which, _dp(a))

    def timestep(self) -> float:
        return lib().tjo_timestep(self.h)

    def step(self, dt: float):
        lib().tjo_step(self.h, dt)

    def run(self, final_time: float, max_steps: int = 1 << 30):
        t = ctypes.c_double(0.0)
        n = lib().tjo_run(self.h, final_time, max_steps, ctypes.byref(t))
        return n, t.value

    def reconstruct(self, d: int):
        """UL, UR of direction d from the current state (MUSCL [+THINC/BVD], p2c, sound speed)."""
        lib().tjo_reconstruct(self.h, d)
        return self.get(self.UL), self.get(self.UR)

    def transverse(self, i: int, j: int = 0, k: int = 0) -> float:
        """transverseWaveSpeed of valid cell (i, j, k) of U1."""
        return lib().tjo_transverse(self.h, i, j, k)

    def face_flux(self, d: int, left: np.ndarray, right: np.ndarray) -> np.ndarray:
        l = np.ascontiguousarray(left, dtype=np.float64)
        r = np.ascontiguousarray(right, dtype=np.float64)
        out = np.zeros(self.ncomp)
        lib().tjo_face_flux(self.h, d, _dp(l), _dp(r), _dp(out))
        return out

    def c2p(self):
        lib().tjo_c2p(self.h)

    def p2c(self):
        lib().tjo_p2c(self.h)

    def sound(self):
        lib().tjo_sound(self.h)

    def cleanup_alpha(self) -> int:
        return lib().tjo_cleanup_alpha(self.h)

    def cleanup_v(self):
        lib().tjo_cleanup_v(self.h)

    def plastic(self):
        lib().tjo_plastic(self.h)

    def reactive(self, dt: float):
        """reactiveUpdate (reactive.cpp:60-103) on U1 with step dt."""
        lib().tjo_reactive(self.h, dt)

    def geometric_source(self, dt: float):
        """geometricSourceTerm (radial.cpp:37-54) on U1 with step dt."""
        lib().tjo_geometric_source(self.h, dt)

    def status(self) -> int:
        return lib().tjo_status(self.h)

    def driver_dt(self, n: int, t: float, final_time: float = float("inf")) -> float:
        """main.cpp:128-143: CFL dt, clamp to the final time, x0.2 for n < 5."""
        dt = self.timestep()
        if t + dt > final_time:
            dt = final_time - t
        if n < 5:
            dt *= 0.2
        return dt


def set_threads(n: int) -> None:
    """OpenMP threads of the oracle's box loops."""
    lib().tjo_set_threads(int(n))
