"""MultiFabs of one FAB per box through the C ABI (tj_fabset_*).

An AMReX MultiFab keeps one ghost-grown FAB per box of its BoxArray.  Here S
and T are such MultiFabs (GpuFabLevel: one float64 device tensor per box of
the level's maxSize decomposition, 2 ghost cells per side, padded rows), and
every entry point runs on them directly.  The results must be:

* bit-exact against the oracle, step by step in FP64, on the kernel families
  the level picks: the 2D stage kernels, the flux-plane and march-fused
  passes, the 1D THINC pass with box edges, the solid 5-wave passes in 2D and
  3D (Us-up, Romenskii and Wilkins), and through the device loop (tj_run) the
  geometric source and the mixture/reactive source;
* bit-identical to the same run on a level array, in the mixed form and
  through a captured graph;
* with every ghost cell and row pad as the caller left it (NaN): the library
  reads and writes each cell in the FAB that owns it and nothing else.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from tests.parity_util import exact_fraction, field_errors, max_rel

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _fab_level(pb, fuse=-1):
    from tenjin_amd import GpuFabLevel

    g = GpuFabLevel(pb, ghost=2, pad=3)
    g.set_tuning(-1, fuse)
    g.init_from_prims(pb.prims, pb.shift_pressure)
    return g


def _ghosts_untouched(g):
    gh = g.ghosts()
    assert gh.size > 0 and np.all(np.isnan(gh)), "a ghost cell or row pad was written"


def _stepwise(pb, nsteps, fuse=-1):
    from oracle.oracle import Oracle

    print(f"\n[case] fabs {pb.name} n={pb.n[:pb.dim]} box={pb.max_grid_size} thinc={pb.THINC} steps={nsteps} "
          f"fuse={fuse}", flush=True)
    o = Oracle(pb)
    g = _fab_level(pb, fuse)
    print(f"[case] {len(g.boxes)} FABs", flush=True)
    assert len(g.boxes) > 1
    for n in range(nsteps):
        dto = o.timestep()
        dtg = g.compute_dt()
        assert dtg == dto, f"step {n}: dt {dtg!r} vs oracle {dto!r}"
        dt = 0.2 * dto if n < 5 else dto
        o.step(dt)
        g.set_dt(dt, reset_clock=1)
        g.step()
        so, sg = o.get(), g.get_state()
        err = max_rel(pb, so, sg)
        assert err <= TOL, f"step {n}: {sorted(field_errors(pb, so, sg).items(), key=lambda kv: -kv[1])[:4]}"
        assert exact_fraction(pb, so, sg) == 1.0, f"step {n}: not bit-exact (max rel {err:.3e})"
    assert g.status() == 0
    _ghosts_untouched(g)
    g.close()


def _tube_thinc():
    from tenjin_amd import configs

    pb = configs.shock_tube_1d(n=300, max_grid_size=40)
    pb.bc_lo = (1, 0, 0)
    pb.bc_hi = (1, 0, 0)
    pb.THINC = 1
    return pb


def _inclusion():
    from tenjin_amd import configs

    pb = configs.shock_inclusion_2d(nx=96, ny=40, box=24)  # boxes of 24 x (16, 24): uneven
    pb.bc_lo = (0, 1, 0)
    pb.bc_hi = (1, 0, 0)
    return pb


@pytest.mark.parametrize("fuse", [2, 1, 0])
def test_fabs_inclusion_2d_stepwise(fuse):
    _stepwise(_inclusion(), 16, fuse=fuse)


def test_fabs_tube_thinc_reflective_stepwise():
    _stepwise(_tube_thinc(), 30)


@pytest.mark.parametrize("dim,n,box,steps", [(2, 48, 24, 10), (3, 20, 10, 5)])
def test_fabs_plate_impact_stepwise(dim, n, box, steps):
    from tenjin_amd import configs

    _stepwise(configs.plate_impact(dim=dim, n=n, box=box), steps)


def test_fabs_plate_impact_romenskii_wilkins_stepwise():
    from tenjin_amd import configs

    _stepwise(configs.plate_impact_rw(dim=2, n=48, box=24), 10)


@pytest.mark.parametrize("kind", ["explosion_rz", "reactive"])
def test_fabs_device_loop_sources(kind):
    """tj_run over FABs on the levels with the geometric source and with the
    mixture's reactive source, against the oracle's loop."""
    from oracle.oracle import Oracle
    from tenjin_amd import configs

    if kind == "explosion_rz":
        pb = configs.underwater_explosion(2, 64, box=24, thinc=1)
    else:
        pb = configs.reactive_tube(2, 64, box=24, thinc=1)
    print(f"\n[case] fabs device loop {pb.name} n={pb.n[:pb.dim]} box={pb.max_grid_size}", flush=True)
    o = Oracle(pb)
    no, to = o.run(float("inf"), max_steps=25)
    g = _fab_level(pb)
    ng, tg = g.run(float("inf"), max_steps=25, check_every=25)
    assert (ng, tg) == (no, to)
    so, sg = o.get(), g.get_state()
    assert max_rel(pb, so, sg) <= TOL
    assert exact_fraction(pb, so, sg) == 1.0
    assert g.status() == 0
    _ghosts_untouched(g)
    g.close()


@pytest.mark.parametrize("kind", ["inclusion", "plate"])
def test_fabs_mixed_and_graph_match_level_array(kind):
    """The mixed form (tj_run) and a captured graph of FP64 loop bodies give
    the same bits on FABs as on a level array."""
    from tenjin_amd import TJ_PREC_MIXED, GpuLevel, configs

    pb = _inclusion() if kind == "inclusion" else configs.plate_impact(dim=2, n=48, box=24)
    print(f"\n[case] fabs vs level array {pb.name}: mixed run, FP64 graph", flush=True)
    g = _fab_level(pb)
    ref = GpuLevel(pb)
    ref.init_from_prims(pb.prims, pb.shift_pressure)
    ng, tg = g.run(float("inf"), max_steps=12, prec=TJ_PREC_MIXED, check_every=4)
    nr, tr = ref.run(float("inf"), max_steps=12, prec=TJ_PREC_MIXED, check_every=4)
    assert (ng, tg) == (nr, tr)
    assert np.array_equal(g.get_state(), ref.get_state())
    for lvl in (g, ref):
        lvl.begin_run(float("inf"))
        lvl.graph_capture(5)
        lvl.graph_launch()
        lvl.graph_launch()
    assert g.clock() == ref.clock()
    assert np.array_equal(g.get_state(), ref.get_state())
    _ghosts_untouched(g)
    g.close()
    ref.close()


def test_fabset_rejects_foreign_boxes_and_mixed_forms():
    from tenjin_amd import Fab, GpuFabLevel, MFView, TenjinError, load_library

    pb = _inclusion()
    g = GpuFabLevel(pb)
    lib = load_library()
    # a FAB list that is not the level's decomposition (the first box moved by one cell)
    recs = (Fab * len(g.boxes))()
    for q, (lo, hi) in enumerate(g.boxes):
        t = g.fabS[q]
        recs[q].data = t.data_ptr()
        for d in range(3):
            recs[q].lo[d], recs[q].hi[d], recs[q].glo[d] = lo[d], hi[d], lo[d] - (2 if d < 2 else 0)
        recs[q].jstride = t.shape[3]
        recs[q].kstride = t.shape[3] * t.shape[2]
        recs[q].nstride = recs[q].kstride * t.shape[1]
    recs[0].hi[0] -= 1
    fs = ctypes.c_void_p()
    assert lib.tj_fabset_create(g.ctx, len(g.boxes), recs, g.ncomp, ctypes.byref(fs)) == -1
    assert b"decomposition" in lib.tj_last_error(g.ctx)
    # one FAB set view and one level array in one call
    import torch

    arr = torch.zeros((g.ncomp,) + pb.shape, dtype=torch.float64, device="cuda")
    v = MFView()
    v.data = arr.data_ptr()
    v.jstride, v.kstride, v.nstride = pb.shape[2], pb.shape[2] * pb.shape[1], arr[0].numel()
    v.ncomp = g.ncomp
    assert lib.tj_step(g.ctx, ctypes.byref(g.vS), ctypes.byref(v), 0) == -1
    assert b"must all be level arrays or all tj_fabsets" in lib.tj_last_error(g.ctx)
    with pytest.raises(TenjinError):
        g.init_from_device(arr)
    g.close()


@pytest.mark.parametrize("fab", [False, True])
def test_box_op_matches_oracle_and_level_ops(fab):
    """tj_box_op box by box: conservativeToPrimitive of a perturbed state and
    the box dt minimum against the oracle's level-wide calls; stress / HJ2 /
    devH, the plastic return and cleanUpV against the level-wide entry points
    (tj_derive, tj_post_step), bit for bit, on a level array and on FABs."""
    from oracle.oracle import Oracle
    from tenjin_amd import (TJ_BOX_C2P, TJ_BOX_DERIVE, TJ_OP_CLEAN_V, TJ_OP_DT, TJ_OP_PLASTIC, GpuFabLevel,
                            GpuLevel, configs)
    from tenjin_amd.configs import TOTAL_E

    pb = configs.plate_impact(dim=2, n=48, box=24)
    print(f"\n[case] box ops {pb.name} fab={fab}", flush=True)
    o = Oracle(pb)
    o.run(float("inf"), max_steps=6)
    s = o.get()
    s[pb.layout.comp(TOTAL_E)] *= 1.0 + 1e-3 * np.sin(np.arange(s[0].size)).reshape(s[0].shape)
    o.set(s)
    o.c2p()
    g = (GpuFabLevel if fab else GpuLevel)(pb)
    g.set_state(s)
    boxes = g.boxes if fab else None
    if boxes is None:
        nb = ctypes.c_int()
        g.lib.tj_level_boxes(g.ctx, ctypes.byref(nb), None)
        lohi = (ctypes.c_int * (6 * nb.value))()
        g.lib.tj_level_boxes(g.ctx, ctypes.byref(nb), lohi)
        boxes = [(tuple(lohi[6 * q:6 * q + 3]), tuple(lohi[6 * q + 3:6 * q + 6])) for q in range(nb.value)]
    assert len(boxes) == 4
    for lo, hi in boxes:
        g.box_op(lo, hi, TJ_BOX_C2P)
    so, sg = o.get(), g.get_state()
    assert exact_fraction(pb, so, sg) == 1.0
    dt = min(g.box_op(lo, hi, TJ_OP_DT) for lo, hi in boxes)
    assert dt == o.timestep()
    # the same level, level-wide
    ref = GpuLevel(pb)
    ref.set_state(g.get_state())
    for lo, hi in boxes:
        g.box_op(lo, hi, TJ_BOX_DERIVE)
    ref.derive()
    assert np.array_equal(g.get_state(), ref.get_state())
    for lo, hi in boxes:
        g.box_op(lo, hi, TJ_OP_PLASTIC | TJ_OP_CLEAN_V)
    ref.post_step(TJ_OP_PLASTIC | TJ_OP_CLEAN_V)
    assert np.array_equal(g.get_state(), ref.get_state())
    if fab:
        _ghosts_untouched(g)
    g.close()
    ref.close()
