"""GPU parity: the gfx950 library against the oracle, through the C ABI.

The FP64 form keeps the reference's operation order, so the bar is bit-exact
equality on the persistent state components (alpha, alphaRho, rho_k, rho,
rhoU, u, E, p) and on dt; the written tolerance 1e-12 (relative max-norm per
component) is the contract's and is asserted as well.
"""
from __future__ import annotations

import numpy as np
import pytest

from tests.parity_util import exact_fraction, field_errors, max_rel

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _pair(pb, seglen=-1, fuse=-1):
    from oracle.oracle import Oracle
    from tenjin_amd import GpuLevel

    o = Oracle(pb)
    g = GpuLevel(pb)
    g.set_tuning(seglen, fuse)
    g.init_from_prims(pb.prims, pb.shift_pressure)
    return o, g


def _stepwise(pb, nsteps, exact=True, seglen=-1, fuse=-1):
    print(f"\n[case] {pb.name} n={pb.n[:pb.dim]} box={pb.max_grid_size} thinc={pb.THINC} "
          f"bc_lo={pb.bc_lo[:pb.dim]} bc_hi={pb.bc_hi[:pb.dim]} steps={nsteps} seglen={seglen} fuse={fuse}",
          flush=True)
    o, g = _pair(pb, seglen, fuse)
    so, sg = o.get(), g.get_state()
    assert max_rel(pb, so, sg) <= TOL
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
        if exact:
            assert exact_fraction(pb, so, sg) == 1.0, f"step {n}: not bit-exact (max rel {err:.3e})"
    assert g.status() == 0
    return o, g


def test_shock_tube_1d_stepwise():
    from tenjin_amd import configs

    _stepwise(configs.shock_tube_1d(), 60)


def test_shock_tube_1d_boxes_reflective():
    from tenjin_amd import configs

    pb = configs.shock_tube_1d(n=300, max_grid_size=40)
    pb.bc_lo = (1, 0, 0)
    pb.bc_hi = (1, 0, 0)
    pb.THINC = 1
    _stepwise(pb, 40)


def test_shock_tube_1d_full_run():
    """The device-resident time loop (tj_run: dt, clamps, steps, NaN stop)
    against the oracle's loop of main.cpp:121-191, to t = 2.4e-4 s."""
    from tenjin_amd import configs

    pb = configs.shock_tube_1d()
    print(f"\n[case] {pb.name} full run to {pb.final_time}", flush=True)
    o, g = _pair(pb)
    n_o, t_o = o.run(pb.final_time)
    n_g, t_g = g.run(pb.final_time, check_every=64)
    assert (n_g, t_g) == (n_o, t_o)
    so, sg = o.get(), g.get_state()
    assert max_rel(pb, so, sg) <= TOL
    assert exact_fraction(pb, so, sg) == 1.0


@pytest.mark.parametrize("thinc,fuse", [(0, 1), (1, 1), (0, 0), (1, 0), (0, 2), (1, 2)])
def test_inclusion_2d_stepwise(thinc, fuse):
    """fuse=2: the stage kernels (x kernel with the x update, y kernel over LDS
    tiles with the rest); fuse=1: flux planes, the update inside the y pass;
    fuse=0: flux planes and a separate update kernel."""
    from tenjin_amd import configs

    pb = configs.shock_inclusion_2d(nx=128, ny=64, box=32, thinc=thinc, shock_x=0.55, drift=(150.0, 60.0))
    _stepwise(pb, 30, fuse=fuse)


@pytest.mark.parametrize("seglen", [5, 16, 64])
def test_inclusion_2d_uneven_boxes(seglen):
    """Box sizes from BoxArray::maxSize that do not divide the domain, THINC on,
    transmissive/reflective mix; marching segments that do not divide ny (5),
    that do (16 does not divide 60 either; 64 covers the column in one)."""
    from tenjin_amd import configs

    pb = configs.shock_inclusion_2d(nx=100, ny=60, box=24, thinc=1, shock_x=0.55, drift=(-120.0, 90.0))
    pb.bc_lo = (1, 0, 0)
    pb.bc_hi = (0, 1, 0)
    _stepwise(pb, 25, seglen=seglen)


@pytest.mark.parametrize("seglen,fuse", [(16, 1), (3, 1), (7, 0)])
def test_inclusion_3d_stepwise(seglen, fuse):
    from tenjin_amd import configs

    _stepwise(configs.shock_inclusion_3d(), 12, seglen=seglen, fuse=fuse)


def test_inclusion_2d_full_size_two_steps():
    """Full-size workload (2048 x 1024, 64^2 boxes, THINC): two steps from the
    initial condition against the oracle."""
    from tenjin_amd import configs

    _stepwise(configs.shock_inclusion_2d(), 2)


def test_inflow_balance_full_size():
    """Size-independent property at the full size (2048 x 1024, 64^2 boxes,
    THINC): over 40 device-resident steps the cell sums of each material's
    mass and of the total energy change exactly by the post-shock inflow
    through the x = 0 boundary (interior fluxes telescope)."""
    from tenjin_amd import GpuLevel, configs
    from tests.test_oracle import inflow_balance

    pb = configs.shock_inclusion_2d()
    print(f"\n[case] inflow balance {pb.n[:2]}", flush=True)
    g = GpuLevel(pb)
    g.init_from_prims(pb.prims)
    s0 = g.get_state()
    g.begin_run()
    g.advance(40)
    s1 = g.get_state()
    t, _, n = g.clock()
    assert n == 40
    for got, expect, total in inflow_balance(pb, s0, s1, t):
        assert abs(got - expect) <= 1e-13 * abs(total)
    assert g.status() == 0


def test_nan_stops_loop():
    """A NaN planted in the state (from the host) raises the NaN status bit and
    stops the device loop before t and n advance (main.cpp:183-187)."""
    from tenjin_amd import TJ_ST_NAN, GpuLevel, configs
    from tenjin_amd.configs import TOTAL_E

    pb = configs.shock_tube_1d(n=200)
    print("\n[case] NaN stop", flush=True)
    g = GpuLevel(pb)
    g.init_from_prims(pb.prims)
    s = g.get_state()
    s[pb.layout.comp(TOTAL_E), 0, 0, 50] = np.nan
    g.set_state(s)
    n, t = g.run(1.0, max_steps=10, check_every=1)
    assert g.status() & TJ_ST_NAN
    assert n == 0 and t == 0.0


@pytest.mark.parametrize("dim,n,box,steps", [(2, 48, 24, 12), (3, 20, 10, 6)])
def test_plate_impact_stepwise(dim, n, box, steps):
    """Elastoplastic path (Us-up EOS, 5-wave flux, radial return, cleanUpV)."""
    from tenjin_amd import configs

    _stepwise(configs.plate_impact(dim=dim, n=n, box=box), steps)


@pytest.mark.parametrize("dim,n,box,steps", [(2, 48, 24, 12), (3, 20, 10, 6)])
def test_plate_impact_romenskii_wilkins_stepwise(dim, n, box, steps):
    """The reference's other solid EOSs (RomenskiiSolidEOS copper flyer,
    WilkinsSolidEOS aluminium target, equationofstate.cpp:403-596): density-
    dependent shear modulus in the stress, the 5-wave solver and the radial
    return, bit-exact per step."""
    from tenjin_amd import configs

    _stepwise(configs.plate_impact_rw(dim=dim, n=n, box=box), steps)


def test_plate_impact_3d_full_size_invariants():
    """256^3 elastoplastic impact, 3 device-resident steps: no status bit,
    det V = 1 after cleanUpV, plastic strain non-negative, mass positive."""
    import torch

    from tenjin_amd import GpuLevel, configs
    from tenjin_amd.configs import ALPHARHO, EPSILON, V_TENSOR

    pb = configs.plate_impact()
    print(f"\n[case] plate impact {pb.n} invariants", flush=True)
    g = GpuLevel(pb)
    g.init_from_prims(pb.prims, pb.shift_pressure)
    g.begin_run()
    g.advance(3)
    g.sync()
    assert g.status() == 0
    L = pb.layout
    v0 = L.start[V_TENSOR]
    V = g.S[v0:v0 + 9].reshape(3, 3, -1).permute(2, 0, 1)
    det = torch.linalg.det(V)
    assert float((det - 1.0).abs().max()) < 1e-12
    assert float(g.S[L.comp(EPSILON, 0)].min()) >= 0.0
    assert float(g.S[L.comp(ALPHARHO, 0)].min()) > 0.0
    _, _, n = g.clock()
    assert n == 3


@pytest.mark.parametrize("fuse", [2, 1, 0])
def test_stage_api_matches_step(fuse):
    """tj_rk_stage(1), tj_rk_stage(2) and tj_post_step (cleanUpAlpha, NaN scan,
    getTimeStep) reproduce tj_step bit for bit, and both match the oracle."""
    import torch

    from tenjin_amd import GpuLevel, configs

    pb = configs.shock_inclusion_2d(nx=96, ny=64, box=32, thinc=1, shock_x=0.55, drift=(80.0, -40.0))
    print(f"\n[case] stage api fuse={fuse}", flush=True)
    a, b = GpuLevel(pb), GpuLevel(pb)
    for g in (a, b):
        g.set_tuning(-1, fuse)
        g.init_from_prims(pb.prims)
    for n in range(8):
        dt = a.compute_dt()
        assert b.compute_dt() == dt
        dt = 0.2 * dt if n < 5 else dt
        for g in (a, b):
            g.set_dt(dt, reset_clock=1)
        a.step()
        b.stage(1)
        b.stage(2)
        b.post_step(4 | 16 | 8)
        a.sync()
        b.sync()
        assert torch.equal(a.S, b.S), f"step {n}"
    assert a.status() == 0 and b.status() == 0


@pytest.mark.parametrize("thinc,prec", [(1, 0), (1, 1), (0, 0), (0, 1)])
def test_stage_kernels_match_flux_planes(thinc, prec):
    """The stage kernels (the default for two-dimensional fluid levels) and the
    flux-plane passes give bit-identical states in both precision forms, on
    uneven boxes (maxSize 24 on 100 x 70), mixed boundaries and several tiles
    of the y kernel, over 30 device-resident loop bodies."""
    import torch

    from tenjin_amd import GpuLevel, configs

    pb = configs.shock_inclusion_2d(nx=100, ny=70, box=24, thinc=thinc, shock_x=0.55, drift=(-120.0, 90.0))
    pb.bc_lo = (1, 0, 0)
    pb.bc_hi = (0, 1, 0)
    print(f"\n[case] stage kernels vs flux planes thinc={thinc} prec={prec}", flush=True)
    a, b = GpuLevel(pb), GpuLevel(pb)
    a.set_tuning(-1, 2)
    b.set_tuning(-1, 0)
    for g in (a, b):
        g.init_from_prims(pb.prims)
        g.begin_run(float("inf"))
    for n in range(3):
        a.advance(10, prec)
        b.advance(10, prec)
        a.sync()
        b.sync()
        assert torch.equal(a.S, b.S), f"after {10 * (n + 1)} steps"
    assert a.status() == 0 and b.status() == 0
    assert a.clock() == b.clock()
