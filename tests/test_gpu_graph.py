"""hipGraph form of the device-resident loop (tj_graph_capture /
tj_graph_launch): every replay, with eager work and host synchronisation
between replays, must reproduce the eager loop bit for bit (state, t, n)."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _replays_match(pb, nsteps, replays, prec=0):
    from tenjin_amd import GpuLevel

    print(f"\n[case] graph {pb.name} n={pb.n[:pb.dim]} {nsteps} steps x {replays} prec={prec}", flush=True)
    e = GpuLevel(pb)
    e.init_from_prims(pb.prims, pb.shift_pressure)
    e.begin_run()
    g = GpuLevel(pb)
    g.init_from_prims(pb.prims, pb.shift_pressure)
    g.begin_run()
    g.graph_capture(nsteps, prec)
    for rep in range(replays):
        e.advance(nsteps, prec)
        se = e.get_state()
        g.graph_launch()
        sg = g.get_state()  # eager work and a host sync between replays
        assert np.array_equal(se, sg, equal_nan=True), f"replay {rep}"
        assert e.clock() == g.clock()
    assert e.status() == 0 and g.status() == 0


def test_graph_inclusion_2d_thinc():
    from tenjin_amd import configs

    _replays_match(configs.shock_inclusion_2d(nx=128, ny=64, box=32, thinc=1, shock_x=0.55), 6, 4)


def test_graph_plate_impact_2d_mixed():
    from tenjin_amd import TJ_PREC_MIXED, configs

    _replays_match(configs.plate_impact(dim=2, n=48, box=24), 4, 3, TJ_PREC_MIXED)


def test_graph_shock_tube_full_run():
    """The 1D shock tube run to t = 2.4e-4 s by graph replays of 64 loop
    bodies (bodies after the end are no-ops on the device) equals the oracle's
    loop (main.cpp:121-191)."""
    from oracle.oracle import Oracle
    from tenjin_amd import GpuLevel, configs
    from tests.parity_util import exact_fraction

    pb = configs.shock_tube_1d()
    print("\n[case] graph shock tube full run", flush=True)
    o = Oracle(pb)
    n_o, t_o = o.run(pb.final_time)
    g = GpuLevel(pb)
    g.init_from_prims(pb.prims, pb.shift_pressure)
    g.begin_run(pb.final_time)
    g.graph_capture(64)
    for _ in range(n_o // 64 + 2):
        g.graph_launch()
    t, _, n = g.clock()
    assert (n, t) == (n_o, t_o)
    assert exact_fraction(pb, o.get(), g.get_state()) == 1.0


def test_graph_refuses_other_history_phase():
    """A THINC level advanced by a single stage is at the other ping-pong phase
    of its history planes: the captured graph must not be replayed there."""
    from tenjin_amd import GpuLevel, TenjinError, configs

    pb = configs.shock_tube_1d(n=200)
    pb.THINC = 1
    g = GpuLevel(pb)
    g.init_from_prims(pb.prims)
    g.begin_run()
    g.graph_capture(2)
    g.graph_launch()
    g.set_dt(1e-9)
    g.stage(1)  # one direction pass in 1D: odd phase
    with pytest.raises(TenjinError):
        g.graph_launch()


def test_prof_only_times_the_named_launches_alone():
    """tj_prof_only: with a name, only that kernel's launches carry events (one
    per loop body for the stage-2 y kernel of a 2D level); with None, every
    launch again.  Timing changes no result."""
    from tenjin_amd import GpuLevel, configs

    pb = configs.shock_inclusion_2d(nx=128, ny=64, box=32)
    print(f"\n[case] prof_only {pb.name} n={pb.n[:pb.dim]}", flush=True)
    g = GpuLevel(pb)
    g.init_from_prims(pb.prims, pb.shift_pressure)
    g.begin_run()
    ref = GpuLevel(pb)
    ref.init_from_prims(pb.prims, pb.shift_pressure)
    ref.begin_run()
    g.prof_enable(True)
    g.prof_only("ystage2_post")
    g.advance(3)
    only = g.prof_report()
    assert set(only) == {"ystage2_post"} and only["ystage2_post"][0] == 3
    g.prof_enable(True)  # clears the accumulators
    g.prof_only(None)
    g.advance(2)
    every = g.prof_report()
    assert {"xstage1", "xstage2", "ystage1", "ystage2_post"} <= set(every)
    assert all(cnt == 2 for cnt, _ in every.values())
    # tj_prof_sample: the first and then every 4th named launch carries events
    g.prof_enable(True)
    g.prof_only("ystage2_post")
    g.prof_sample(4)
    g.advance(9)
    sampled = g.prof_report()
    g.prof_sample(1)
    g.prof_only(None)
    g.prof_enable(False)
    assert set(sampled) == {"ystage2_post"} and sampled["ystage2_post"][0] == 3  # launches 1, 5, 9
    with pytest.raises(RuntimeError):
        g.prof_sample(0)
    ref.advance(14)
    assert np.array_equal(g.get_state(), ref.get_state()) and g.clock() == ref.clock()
    g.close()
    ref.close()
