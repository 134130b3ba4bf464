"""Reference-format settings decks (libconfig syntax, initialise.cpp:125-385)
-> Problem, with the pluggable initial-condition geometry."""
from __future__ import annotations

import numpy as np
import pytest

from tenjin_amd import settings
from tenjin_amd.configs import ALPHA, P, RHO_K, VELOCITY

DECK = """
// 1D two-material stiffened-gas shock tube, in the reference's keys
finalTime = 2.4e-4;
numberOfxCells = 1000;  numberOfyCells = 1;  numberOfzCells = 1;
xdomainLength = 1.0;  ydomainLength = 1.0;  zdomainLength = 1.0;
CFL = 0.5;
materials = 2;  mixtures = 0;  states = 2;
interface = 0.7;
numberOfPictures = 10;
SOLID = 0; THINC = 0; RADIAL = 0; PLASTIC = 0; MUSCL = 1; REACTIVE = 0;
THINCbeta = 1.6;
lowBoundary = [0, 0, 0];
highBoundary = [0, 0, 0];
/* materials: water-like stiffened gas, then air */
listOfMaterials = (
  { material = "fluid"; mixture = 0; EOS = "MieGruneisen";
    adiabaticIndex = 4.4; pref = -2.64e9; eref = 0.0; CV = 1.0; },
  { material = "fluid"; mixture = 0; EOS = "Mie" "Gruneisen";   # adjacent strings concatenate
    adiabaticIndex = 1.4; pref = 0.0; eref = 0.0; CV = 1.0; }
);
listOfStates = (
  { material = ( { alpha = 0.999999; rho = 1000.0; }, { alpha = 0.000001; rho = 50.0; } );
    p = 1.0e9; u = 0.0; v = 0.0; w = 0.0; },
  { material = ( { alpha = 0.000001; rho = 1000.0; }, { alpha = 0.999999; rho = 50.0; } );
    p = 1.0e5; u = 0.0; v = 0.0; w = 0.0; }
);
"""


def test_parse_libconfig_subset():
    cfg = settings.parse(DECK)
    assert cfg["numberOfxCells"] == 1000 and cfg["finalTime"] == 2.4e-4 and cfg["lowBoundary"] == [0, 0, 0]
    assert cfg["listOfMaterials"][1]["EOS"] == "MieGruneisen"
    assert cfg["listOfStates"][0]["material"][0]["rho"] == 1000.0
    assert settings.parse('a : true; b = "x\\"y"; c = 0x1F; d = {e = [1.5, -2e3];};') == {
        "a": True, "b": 'x"y', "c": 31, "d": {"e": [1.5, -2e3]}}


def test_deck_gives_the_shock_tube():
    """The deck above is the 1D shock tube: the same materials, the same
    per-cell states by the rule of initialise.cpp:393-402, and, with the
    reference's pressure shift of initialise.cpp:104, the same first steps
    in the oracle as the hand-built problem with that shift."""
    from oracle.oracle import Oracle
    from tenjin_amd import configs

    pb = settings.problem_from_settings(DECK, max_grid_size=1 << 30)
    ref = configs.shock_tube_1d()
    assert pb.dim == 1 and pb.n == ref.n and pb.final_time == ref.final_time and pb.shift_pressure == 1
    for a, b in zip(pb.materials, ref.materials):
        assert (a.gamma, a.pref, a.eref) == (b.gamma, b.pref, b.eref)
    L = pb.layout
    for c in (L.comp(RHO_K, 0), L.comp(RHO_K, 1), L.comp(P), L.comp(VELOCITY, 0, 0)):
        assert np.array_equal(pb.prims[c], ref.prims[c])
    for m in range(2):  # the deck's literals vs 1 - (1 - 1e-6) of the hand-built problem: rounding only
        c = L.comp(ALPHA, m)
        assert np.allclose(pb.prims[c], ref.prims[c], rtol=0, atol=1e-16)
        ref.prims[c] = pb.prims[c]
    ref.shift_pressure = 1
    o1, o2 = Oracle(pb), Oracle(ref)
    for n in range(3):
        dt = o1.driver_dt(n, 0.0)
        assert dt == o2.driver_dt(n, 0.0)
        o1.step(dt)
        o2.step(dt)
    assert np.array_equal(o1.get(), o2.get())


def test_custom_geometry_and_rejections():
    cfg = settings.parse(DECK)
    cfg["numberOfyCells"] = 8
    cfg["numberOfxCells"] = 16
    pb = settings.problem_from_settings(cfg, geometry=lambda i, j, k, pb, cfg: (j >= 4).astype(int))
    assert pb.dim == 2
    assert pb.prims[pb.layout.comp(P)][0, 0, 0] == 1e9 and pb.prims[pb.layout.comp(P)][0, 7, 0] == 1e5
    bad = settings.parse(DECK)
    bad["listOfMaterials"][0]["EOS"] = "RomenskiiSolid"  # a solid EOS on a fluid material
    with pytest.raises(ValueError):
        settings.problem_from_settings(bad)
    bad = settings.parse(DECK)
    bad["listOfMaterials"][0]["EOS"] = "Tillotson"  # not an EOS of the reference
    with pytest.raises(ValueError):
        settings.problem_from_settings(bad)
    bad = settings.parse(DECK)
    bad["mixtures"] = 2  # every mixture would share mixtureIndex 0
    with pytest.raises(ValueError):
        settings.problem_from_settings(bad)
    bad = settings.parse(DECK)
    bad["listOfMaterials"][1]["mixture"] = 1  # a mixture needs the products' *mix constants
    with pytest.raises(KeyError):
        settings.problem_from_settings(bad)


SOLID_DECK = """
numberOfxCells = 8;  numberOfyCells = 1;  numberOfzCells = 1;
xdomainLength = 1.0;
CFL = 0.5;  materials = 2;  states = 2;  interface = 0.5;
SOLID = 1; THINC = 0; PLASTIC = 1; MUSCL = 1;
listOfMaterials = (
  { material = "solid"; EOS = "UsUpSolid"; rho0 = 8930.0; Gamma0 = 2.0; c0 = 3940.0; s = 1.49;
    G0 = 45.0e9; yieldStress = 90.0e6; },
  { material = "solid"; EOS = "UsUpSolid"; rho0 = 2703.0; Gamma0 = 2.0; c0 = 5328.0; s = 1.338;
    G0 = 26.0e9; yieldStress = 300.0e6; }
);
listOfStates = (
  { material = ( { alpha = 0.999999; F = [1.02, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0]; },
                 { alpha = 0.000001; F = [0.95, 0.01, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0]; } );
    p = 0.0; u = 100.0; v = 0.0; w = 0.0; },
  { material = ( { alpha = 0.000001; F = [1.0, 0.0, 0.0, 0.0, 1.1, 0.0, 0.0, 0.0, 1.0]; },
                 { alpha = 0.999999; F = [1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.9]; } );
    p = 0.0; u = 0.0; v = 0.0; w = 0.0; }
);
"""


def test_solid_states_take_the_last_solid_F():
    """getState keeps one F per state and each solid material overwrites it
    (initialise.cpp:237-263): the last solid's F sets V and the density of
    every solid of the state (initialise.cpp:27-46, 74-77)."""
    from tenjin_amd.configs import V_TENSOR

    pb = settings.problem_from_settings(SOLID_DECK)
    L, U = pb.layout, pb.prims
    lastF = [np.array([0.95, 0.01, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0]).reshape(3, 3),
             np.diag([1.0, 1.0, 0.9])]
    rho0 = [8930.0, 2703.0]
    for s, i in ((0, 0), (1, 7)):
        F = lastF[s]
        for m in range(2):
            assert U[L.comp(RHO_K, m)][0, 0, i] == rho0[m] / np.linalg.det(F)
        V = np.array([[U[L.comp(V_TENSOR, 0, r, c)][0, 0, i] for c in range(3)] for r in range(3)])
        # V = sqrt(F F^T) normalised to det 1 (closed form check: V V^T = B / det(B)^(1/3))
        B = F @ F.T
        assert np.allclose(V @ V.T, B / np.linalg.det(B) ** (1.0 / 3.0), rtol=0, atol=1e-14)
        assert np.allclose(V, V.T, atol=1e-15)


ROMENSKII_WILKINS_DECK = """
numberOfxCells = 8;  numberOfyCells = 1;  numberOfzCells = 1;
CFL = 0.5;  materials = 2;  states = 2;  interface = 0.5;
SOLID = 1; THINC = 0; PLASTIC = 1; MUSCL = 1;
listOfMaterials = (
  { material = "solid"; mixture = 0; EOS = "RomenskiiSolid"; adiabaticIndex = 3.0; pref = 0.0; eref = 0.0; CV = 390.0;
    rho0 = 8930.0; K0 = 1.36e11; EOSalpha = 1.3; EOSbeta = 2.2; G0 = 4.5e10; yieldStress = 9.0e7; },
  { material = "solid"; mixture = 0; EOS = "WilkinsSolid"; adiabaticIndex = 3.0; pref = 0.0; eref = 0.0; CV = 900.0;
    rho0 = 2703.0; e1 = 3.8e10; e2 = 3.8e10; e3 = -6.46e10; e4 = -1.14e10; e5 = 2.28e10; G0 = 2.6e10;
    yieldStress = 3.0e8; }
);
listOfStates = (
  { material = ( { alpha = 0.999999; F = [1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0]; },
                 { alpha = 0.000001; F = [1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0]; } );
    p = 0.0; u = 400.0; v = 0.0; w = 0.0; },
  { material = ( { alpha = 0.000001; F = [1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0]; },
                 { alpha = 0.999999; F = [1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0]; } );
    p = 0.0; u = 0.0; v = 0.0; w = 0.0; }
);
"""


def test_romenskii_and_wilkins_decks():
    """The reference's keys for RomenskiiSolid and WilkinsSolid
    (initialise.cpp:149-172) load into EOS kinds 2 and 3 with every constant."""
    pb = settings.problem_from_settings(ROMENSKII_WILKINS_DECK)
    r, w = pb.materials
    assert (r.eos, r.solid, r.plastic, r.gamma, r.CV, r.rho0, r.K0, r.eos_alpha, r.eos_beta, r.G0, r.yield_stress) == (
        2, 1, 1, 3.0, 390.0, 8930.0, 1.36e11, 1.3, 2.2, 4.5e10, 9.0e7)
    assert (w.eos, w.rho0, w.e1, w.e2, w.e3, w.e4, w.e5, w.G0, w.yield_stress) == (
        3, 2703.0, 3.8e10, 3.8e10, -6.46e10, -1.14e10, 2.28e10, 2.6e10, 3.0e8)
    assert pb.shift_pressure == 1 and pb.SOLID == 1 and pb.PLASTIC == 1
