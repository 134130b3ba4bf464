"""The reference's C++ hot-path interface (tenjin_host.h: advance,
TimeStep::getTimeStep, PlasticEOS::plasticUpdate, CellArray clean-ups and
contains_nan, over amrex_lite types) driving the time loop of main.cpp:121-191
from C++ (tests/host/main_loop.cpp), against the oracle's loop: same number of
steps, same final time, bit-identical state."""
from __future__ import annotations

import os

import numpy as np
import pytest

from tests.parity_util import exact_fraction, max_rel

TOL = 1e-12


def test_driver_and_host_library_built():
    from tests.host_driver import DRIVER

    assert os.path.exists(DRIVER)
    assert os.path.exists(os.path.join(os.path.dirname(DRIVER), "..", "..", "tenjin_amd", "libtenjin_host.so"))


def _check(pb, final_time, max_steps, tmp_path, **kw):
    from oracle.oracle import Oracle
    from tests import host_driver

    print(f"\n[case] host interface {pb.name} n={pb.n[:pb.dim]} thinc={pb.THINC} solid={pb.SOLID} "
          f"final={final_time} max_steps={max_steps} {kw}", flush=True)
    o = Oracle(pb)
    n_o, t_o = o.run(final_time, max_steps)
    t_g, n_g, nan_stop, sg = host_driver.run(pb, final_time, max_steps, str(tmp_path), **kw)
    assert nan_stop == 0
    assert (n_g, t_g) == (n_o, t_o)
    so = o.get()
    assert max_rel(pb, so, sg) <= TOL
    assert exact_fraction(pb, so, sg) == 1.0


@pytest.mark.gpu
def test_host_interface_shock_tube_full_run(tmp_path):
    from tenjin_amd import configs

    pb = configs.shock_tube_1d()
    _check(pb, pb.final_time, 1 << 30, tmp_path)


@pytest.mark.gpu
def test_host_interface_inclusion_2d_thinc(tmp_path):
    from tenjin_amd import configs

    pb = configs.shock_inclusion_2d(nx=128, ny=64, box=32, thinc=1, shock_x=0.55, drift=(150.0, 60.0))
    _check(pb, 1.0, 25, tmp_path)


@pytest.mark.gpu
def test_host_interface_plate_impact_2d(tmp_path):
    from tenjin_amd import configs

    pb = configs.plate_impact(dim=2, n=48, box=24)
    _check(pb, 1.0, 10, tmp_path)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["tube_thinc", "inclusion", "plate2d", "plate3d_rw"])
@pytest.mark.parametrize("fab_per_box", [True, False])
def test_host_interface_per_box_tier(tmp_path, kind, fab_per_box):
    """The reference's loop on CellArrays of one ghost-grown FArrayBox per box
    (AMReX's MultiFab storage; the library works on the FABs in place) or of
    one level array, bit-exact against the oracle; then the per-box tier
    (BoxAccessCellArray per-box methods, boxPlasticUpdate, boxTimeStepFunction)
    box by box against the level-wide calls, bit for bit (the driver's
    --box-checks)."""
    from tenjin_amd import configs

    if kind == "tube_thinc":
        pb = configs.shock_tube_1d(n=300, max_grid_size=40)
        pb.THINC = 1
        steps = 30
    elif kind == "inclusion":
        pb = configs.shock_inclusion_2d(nx=96, ny=40, box=24, thinc=1, shock_x=0.55, drift=(150.0, 60.0))
        steps = 15
    elif kind == "plate2d":
        pb = configs.plate_impact(dim=2, n=48, box=24)
        steps = 8
    else:
        pb = configs.plate_impact_rw(dim=3, n=20, box=10)
        steps = 4
    _check(pb, 1.0, steps, tmp_path, fab_per_box=fab_per_box, box_checks=True)


@pytest.mark.gpu
def test_host_interface_plotfiles(tmp_path):
    """WriteSingleLevelPlotfile of the C++ interface (main.cpp:104-107,
    195-199): the initial and final plotfiles hold every component of U1 with
    AccessPattern's names, the final one equal to the oracle's state."""
    from oracle.oracle import Oracle
    from tenjin_amd import configs
    from tenjin_amd.plotfile import read_plotfile
    from tests import host_driver

    pb = configs.shock_inclusion_2d(nx=96, ny=64, box=32, thinc=1, shock_x=0.55)
    print("\n[case] host interface plotfiles", flush=True)
    o = Oracle(pb)
    n_o, t_o = o.run(1.0, 12)
    root = str(tmp_path / "plt")
    t_g, n_g, _, sg = host_driver.run(pb, 1.0, 12, str(tmp_path), plotroot=root)
    meta0, d0 = read_plotfile(root + "00000")
    meta, d = read_plotfile(root + f"{n_g:05d}")
    assert meta["step"] == n_g == n_o and meta["time"] == t_g == t_o and meta0["time"] == 0.0
    assert meta["names"][0] == "alpha0_00" and meta["names"][pb.layout.comp(2)] == "rho_k0_00"
    assert len(meta["boxes"]) == 3 * 2 and meta["dim"] == 2
    assert np.array_equal(d, sg[: d.shape[0]], equal_nan=True)
    assert exact_fraction(pb, o.get(), d) == 1.0
