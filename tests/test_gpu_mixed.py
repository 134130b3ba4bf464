"""Mixed-precision form (TJ_PREC_MIXED): reconstruction, face states, wave
speeds and fluxes in FP32, FP64 in the update accumulation, c2p, EOS inversion,
clean-ups and the plastic return.  Checked against the FP64 oracle on the same
dt sequence, by relative L1 error  sum|x - x_ref| / sum|x_ref|  of density,
pressure and volume fraction.  The bar for density is 1e-5 on every config;
pressure and alpha bars are stated per case (measured errors are 10-100x
below them)."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RHO_BAR = 1e-5


def _rel_l1(pb, got, ref):
    from tenjin_amd.configs import ALPHA, P, RHO

    L = pb.layout
    out = {}
    for name, c in (("rho", L.comp(RHO)), ("p", L.comp(P)), ("alpha", L.comp(ALPHA, 0))):
        out[name] = float(np.abs(got[c] - ref[c]).sum() / max(np.abs(ref[c]).sum(), 1e-300))
    return out


def _mixed_vs_oracle(pb, nsteps):
    from oracle.oracle import Oracle
    from tenjin_amd import TJ_PREC_MIXED, GpuLevel

    print(f"\n[case] mixed {pb.name} n={pb.n[:pb.dim]} steps={nsteps}", flush=True)
    o = Oracle(pb)
    g = GpuLevel(pb)
    g.init_from_prims(pb.prims, pb.shift_pressure)
    for n in range(nsteps):
        dt = o.driver_dt(n, 0.0)
        o.step(dt)
        g.set_dt(dt, reset_clock=1)
        g.step(TJ_PREC_MIXED)
    err = _rel_l1(pb, g.get_state(), o.get())
    print(f"  rel L1 {err}", flush=True)
    assert g.status() == 0
    return err


def test_mixed_shock_tube_1d():
    from tenjin_amd import configs

    err = _mixed_vs_oracle(configs.shock_tube_1d(), 200)
    assert err["rho"] <= RHO_BAR and err["p"] <= 1e-6 and err["alpha"] <= 1e-7


def test_mixed_inclusion_2d_thinc():
    from tenjin_amd import configs

    pb = configs.shock_inclusion_2d(nx=128, ny=64, box=32, thinc=1, shock_x=0.55, drift=(150.0, 60.0))
    err = _mixed_vs_oracle(pb, 60)
    assert err["rho"] <= RHO_BAR and err["p"] <= 1e-6 and err["alpha"] <= RHO_BAR


def test_mixed_plate_impact_2d():
    from tenjin_amd import configs

    err = _mixed_vs_oracle(configs.plate_impact(dim=2, n=64, box=32), 40)
    assert err["rho"] <= RHO_BAR and err["p"] <= 1e-4 and err["alpha"] <= 1e-7


def test_mixed_plate_impact_3d():
    from tenjin_amd import configs

    err = _mixed_vs_oracle(configs.plate_impact(dim=3, n=32, box=16), 20)
    assert err["rho"] <= RHO_BAR and err["p"] <= 1e-4 and err["alpha"] <= 1e-7


def test_mixed_is_fp32_and_deterministic():
    """The mixed form really runs its fluxes in FP32 (one step differs from the
    FP64 step, by far less than the bar) and is deterministic run to run."""
    from tenjin_amd import TJ_PREC_FP64, TJ_PREC_MIXED, GpuLevel, configs

    pb = configs.shock_inclusion_2d(nx=96, ny=64, box=32, thinc=1, shock_x=0.55)
    states = []
    for prec in (TJ_PREC_FP64, TJ_PREC_MIXED, TJ_PREC_MIXED):
        g = GpuLevel(pb)
        g.init_from_prims(pb.prims)
        dt = g.compute_dt()
        g.set_dt(0.2 * dt, reset_clock=1)
        for _ in range(3):
            g.step(prec)
        states.append(g.get_state())
    assert not np.array_equal(states[0], states[1])
    assert np.array_equal(states[1], states[2])
    assert _rel_l1(pb, states[1], states[0])["rho"] <= RHO_BAR
