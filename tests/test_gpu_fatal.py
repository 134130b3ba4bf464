"""The NaN and fatal paths of the time loop against the oracle's.

The reference stops its loop (main.cpp:183-187) when the state holds a NaN
(CellArray::contains_nan, cellarray.cpp:101-117), and exits on "too many Nans"
or a non-positive alpha sum in cleanUpAlpha (boxaccesscellarray.cpp:530-599)
and on NaN wave speeds in the 5-wave solver (hllc.cpp:169-200, 367-373).  The
device loop raises the TJ_ST_* bit of each condition and stops at the same
step.  Each case plants the same bad cell in the oracle's state and in the
GPU's and runs both loops.  The fatal bits, the step count and the time must
agree.  On a clean NaN stop (no fatal bit) the NaN bit and the state must
agree too: NaNs at the same places, every other value bit-equal.  After a
fatal condition the reference has exited, so its state is not compared.
The planted values are data (NaN, zero volume fractions): no index in any
kernel derives from them.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _plant(pb, s, what, cell):
    from tenjin_amd.configs import ALPHA, P, TOTAL_E

    L = pb.layout
    if what == "nan_E":
        s[(L.comp(TOTAL_E),) + cell] = np.nan
    elif what == "nan_alpha1":  # one NaN alpha: cleanUpAlpha's Nan-index quirk repairs it
        s[(L.comp(ALPHA, 1),) + cell] = np.nan
    elif what == "nan_alpha2":  # "too many Nans"
        s[(L.comp(ALPHA, 0),) + cell] = np.nan
        s[(L.comp(ALPHA, 1),) + cell] = np.nan
    elif what == "zero_alpha":  # total alpha <= 0
        s[(L.comp(ALPHA, 0),) + cell] = 0.0
        s[(L.comp(ALPHA, 1),) + cell] = 0.0
    elif what == "nan_p":
        s[(L.comp(P),) + cell] = np.nan
    else:
        raise ValueError(what)


def _problem(kind):
    from tenjin_amd import configs

    if kind == "tube":
        return configs.shock_tube_1d(n=200), (0, 0, 120)
    if kind == "tube_thinc":
        pb = configs.shock_tube_1d(n=200, max_grid_size=50)
        pb.THINC = 1
        return pb, (0, 0, 138)
    if kind == "inclusion":
        return configs.shock_inclusion_2d(nx=64, ny=32, box=16), (0, 16, 26)
    if kind == "plate":
        return configs.plate_impact(dim=2, n=32, box=16), (0, 16, 14)
    raise ValueError(kind)


CASES = [
    ("tube", "nan_E"),
    ("tube", "nan_alpha1"),
    ("tube", "nan_alpha2"),
    ("tube", "zero_alpha"),
    ("tube_thinc", "nan_alpha1"),
    ("tube_thinc", "nan_alpha2"),
    ("inclusion", "nan_alpha2"),
    ("inclusion", "zero_alpha"),
    ("plate", "nan_p"),
    ("plate", "nan_alpha2"),
]


@pytest.mark.parametrize("kind,what", CASES)
def test_fatal_paths_match_oracle(kind, what):
    from oracle.oracle import Oracle
    from tenjin_amd import GpuLevel

    pb, cell = _problem(kind)
    print(f"\n[case] fatal path {kind} {what} cell={cell}", flush=True)
    o = Oracle(pb)
    s = o.get()
    _plant(pb, s, what, cell)
    o.set(s)
    no, to = o.run(1.0, max_steps=6)
    so, sto = o.get(), o.status()

    g = GpuLevel(pb)
    g.init_from_prims(pb.prims)
    s = g.get_state()
    _plant(pb, s, what, cell)
    g.set_state(s)
    ng, tg = g.run(1.0, max_steps=6, check_every=1)
    sg, stg = g.get_state(), g.status()
    g.close()

    assert sto != 0, "the planted cell must trip the reference's checks"
    assert (stg & ~1) == (sto & ~1), f"fatal bits {stg:#x} vs oracle {sto:#x}"
    assert (ng, tg) == (no, to)
    if sto & ~1:
        return
    assert stg == sto, f"status {stg:#x} vs oracle {sto:#x}"
    comps = pb.layout.state_comps()
    a, b = sg[comps], so[comps]
    assert np.array_equal(np.isnan(a), np.isnan(b)), "NaNs at different places"
    fin = ~np.isnan(b)
    assert np.array_equal(a[fin], b[fin]), "finite values differ"
