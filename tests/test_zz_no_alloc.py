"""GPU test of tj_run_static, the time loop that allocates nothing (its
buffers live in the library's file-scope device array): through the C ABI,
two calls in a row -- a 2D shock/inclusion with THINC on uneven boxes, then
a smaller 1D shock tube -- and a third repeating the first in the other
precision, each bit-identical to tj_bind_level + tj_run on the same initial
state; and a level beyond the array's size refused with TJ_ESCRATCH."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TJ_ESCRATCH = -5


def _static_lib():
    from tenjin_amd import MFView, LevelDesc, load_library

    L = load_library()
    mfp = ctypes.POINTER(MFView)
    L.tj_run_static.argtypes = [ctypes.POINTER(LevelDesc), mfp, mfp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_int)]
    return L


def _run_both(pb, steps, prec):
    """tj_run on a bound level, and tj_run_static from the same initial state."""
    import torch

    from tenjin_amd import GpuLevel, level_desc

    L = _static_lib()
    g = GpuLevel(pb)
    g.init_from_prims(pb.prims, pb.shift_pressure)
    S0 = g.S.clone()
    n_ref, t_ref = g.run(float("inf"), max_steps=steps, prec=prec, check_every=8)
    ref = g.S.clone()
    g.close()
    S = S0.clone()
    T = torch.zeros_like(S)
    torch.cuda.synchronize()
    vS, vT = g._view(S), g._view(T)
    d = level_desc(pb)
    t = ctypes.c_double()
    n = ctypes.c_int()
    rc = L.tj_run_static(ctypes.byref(d), ctypes.byref(vS), ctypes.byref(vT), float("inf"), steps, prec, 8, None,
                         ctypes.byref(t), ctypes.byref(n))
    torch.cuda.synchronize()
    assert rc == 0
    assert n.value == n_ref == steps and t.value == t_ref
    assert torch.equal(S, ref)


def test_run_static_matches_run_twice_then_smaller():
    from tenjin_amd import TJ_PREC_FP64, TJ_PREC_MIXED, configs

    big = configs.shock_inclusion_2d(nx=96, ny=64, box=24)
    small = configs.shock_tube_1d(n=400)
    print("\n[case] tj_run_static: 2D inclusion 96x64, THINC, boxes of 24 (uneven in y), FP64", flush=True)
    _run_both(big, 30, TJ_PREC_FP64)
    print("[case] tj_run_static: 1D tube 400 cells (smaller), mixed", flush=True)
    _run_both(small, 40, TJ_PREC_MIXED)
    print("[case] tj_run_static: 2D inclusion again, mixed (history cleared)", flush=True)
    _run_both(big, 30, TJ_PREC_MIXED)


def test_run_static_refuses_levels_beyond_its_scratch():
    import torch

    from tenjin_amd import MFView, TJ_PREC_FP64, configs, level_desc

    L = _static_lib()
    d = level_desc(configs.shock_inclusion_2d(nx=96, ny=64, box=24))
    d.n[0], d.n[1] = 2048, 1024  # 2D THINC: about 256 bytes of buffers per cell, 0.5 GB
    a = torch.zeros(16, dtype=torch.float64, device="cuda")
    b = torch.zeros(16, dtype=torch.float64, device="cuda")
    vs, vt = MFView(), MFView()
    vs.data, vt.data = a.data_ptr(), b.data_ptr()
    t = ctypes.c_double()
    n = ctypes.c_int()
    rc = L.tj_run_static(ctypes.byref(d), ctypes.byref(vs), ctypes.byref(vt), 1.0, 1, TJ_PREC_FP64, 1, None,
                         ctypes.byref(t), ctypes.byref(n))
    assert rc == TJ_ESCRATCH
    assert np.count_nonzero(a.cpu().numpy()) == 0
