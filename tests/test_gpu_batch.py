"""Batched ensemble (tj_batch_*): independent 1D and 2D fluid problems of
uneven size, each with its own materials, states and time step, advanced
together; every problem must equal the same problem run alone through the
single-level device loop, bit for bit (state, t, n), in both precision forms,
on the stage kernels with any y-tile shape and on the flux-plane passes with
any marching segment length, and through graph replays."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _small_ensemble(thinc=1):
    from tenjin_amd import configs

    return configs.ensemble(seed=11, n1d=3, n2d=3, n1d_range=(200, 700), side_range=(40, 90), thinc=thinc)


def _alone(pb, nsteps, prec):
    from tenjin_amd import GpuLevel

    g = GpuLevel(pb)
    g.init_from_prims(pb.prims, pb.shift_pressure)
    g.begin_run()
    g.advance(nsteps, prec)
    t, _, n = g.clock()
    st = g.status()
    return g.get_state(), t, n, st


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_matches_single_levels(prec):
    from tenjin_amd import GpuBatch

    probs = _small_ensemble()
    print(f"\n[case] batch of {len(probs)} prec={prec} sizes={[p.n[:p.dim] for p in probs]}", flush=True)
    b = GpuBatch(probs)
    b.init_from_prims()
    b.begin_run()
    nsteps = 25
    b.advance(nsteps, prec)
    t, n, st = b.clocks()
    for p, pb in enumerate(probs):
        s1, t1, n1, st1 = _alone(pb, nsteps, prec)
        assert (t[p], n[p], st[p]) == (t1, n1, st1), f"problem {p}"
        assert np.array_equal(b.state(p), s1, equal_nan=True), f"problem {p}"


def test_batch_seglen_and_graph_replays():
    """The flux-plane form with per-problem segment lengths against the stage
    kernels (the default for the 2D problems): the schedule changes only;
    graph replays of the batch equal eager batch steps."""
    from tenjin_amd import GpuBatch

    probs = _small_ensemble()
    print("\n[case] batch flux planes with seglens + graph vs stage kernels", flush=True)
    e = GpuBatch(probs)
    e.init_from_prims()
    e.begin_run()
    g = GpuBatch(probs)
    g.set_stage2d(False)
    for p in range(len(probs)):
        g.set_seglen(p, 3 + 5 * p)
    g.init_from_prims()
    g.begin_run()
    g.graph_capture(4)
    for rep in range(3):
        e.advance(4)
        g.graph_launch()
        for p in range(len(probs)):
            assert np.array_equal(e.state(p), g.state(p), equal_nan=True), f"replay {rep} problem {p}"
        assert e.clocks() == g.clocks()


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_tile_shapes_and_graph_replays(prec):
    """Every y-tile shape of the stage kernels, mixed over the 2D problems,
    gives the default's bits; graph replays equal eager steps."""
    from tenjin_amd import BATCH_TILES, GpuBatch

    probs = _small_ensemble()
    print(f"\n[case] batch tile shapes {BATCH_TILES} prec={prec} + graph", flush=True)
    e = GpuBatch(probs)
    e.init_from_prims()
    e.begin_run()
    g = GpuBatch(probs)
    for p in range(len(probs)):
        g.set_tile(p, (p + 1) % len(BATCH_TILES))
    g.init_from_prims()
    g.begin_run()
    g.graph_capture(5, prec)
    for rep in range(3):
        e.advance(5, prec)
        g.graph_launch()
        for p in range(len(probs)):
            assert np.array_equal(e.state(p), g.state(p), equal_nan=True), f"replay {rep} problem {p}"
        assert e.clocks() == g.clocks()


def test_level_seglen_tuner():
    """The level tuner times its candidates in turns on a scratch level and
    returns one of them; a level run with the chosen segment length equals a
    run with the library's own pick, bit for bit."""
    from tenjin_amd import autotune, configs

    pb = configs.shock_inclusion_2d(nx=96, ny=48, box=32)
    print("\n[case] level seglen tuner 96x48", flush=True)
    sl, timing = autotune.tune_level_seglen(pb, 0, at_step=4, steps=2, rounds=2)
    assert set(timing) == set(autotune.LEVEL_CANDIDATES) and sl in timing
    assert all(t > 0.0 for t in timing.values())

    def run(seglen):
        from tenjin_amd import GpuLevel

        g = GpuLevel(pb)
        g.set_tuning(seglen, 0)
        g.init_from_prims(pb.prims, pb.shift_pressure)
        g.begin_run()
        g.advance(6)
        out = g.get_state(), g.clock()
        g.close()
        return out

    (s0, c0), (s1, c1) = run(0), run(max(sl, 8))
    assert c0 == c1 and np.array_equal(s0, s1, equal_nan=True)
