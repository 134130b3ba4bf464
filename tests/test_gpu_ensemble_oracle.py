"""The batched ensemble (tj_batch_*) against the oracle: problems drawn by the
ensemble's own generator (stiffened-gas and Mie-Gruneisen pairs with seeded
gamma, p_inf, reference energies, density and pressure ratios; 1D and 2D
tubes of uneven size; MUSCL + THINC + HLLC), each advanced with its own time
step inside one batch:

* FP64: every problem's state, t and n equal the oracle's own time loop
  (main.cpp:121-191) bit for bit;
* mixed form: every problem's density within 1e-5 relative L1 of the oracle
  after the same loop bodies;
* the full 256-problem ensemble: four of its problems (the largest and the
  smallest of each dimension) bit-exact against the oracle over 20 FP64 steps;
  then 200 steps of each form in which every problem finishes cleanly and
  keeps each material's mass to rounding while no mass crosses its boundary
  (the normal velocity of every boundary cell still exactly 0).  In the mixed
  form this holds because the HLLC star factor is divided correctly rounded:
  with the 1-ulp FP32 reciprocal a quiet region carries a flux S_K (mult - 1) U,
  and two quiet regions' different such fluxes moved about 6e-9 of a material's
  mass per step in and out through the open ends.
"""
from __future__ import annotations

import numpy as np
import pytest

from tests.parity_util import exact_fraction, max_rel

pytestmark = pytest.mark.gpu


def _oracle_run(pb, steps):
    from oracle.oracle import Oracle

    o = Oracle(pb)
    n, t = o.run(float("inf"), max_steps=steps)
    return o.get(), t, n


def _rel_l1(a, b):
    return float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))


@pytest.mark.parametrize("prec", [0, 1])
def test_ensemble_generator_against_oracle(prec):
    from tenjin_amd import GpuBatch, configs

    probs = configs.ensemble(seed=12, n1d=4, n2d=4, n1d_range=(200, 600), side_range=(40, 72))
    steps = 30
    print(f"\n[case] ensemble vs oracle prec={prec} sizes={[p.n[:p.dim] for p in probs]}", flush=True)
    b = GpuBatch(probs)
    b.init_from_prims()
    b.begin_run()
    b.advance(steps, prec)
    t, n, st = b.clocks()
    for p, pb in enumerate(probs):
        so, to, no = _oracle_run(pb, steps)
        sg = b.state(p)
        assert st[p] == 0 and n[p] == no == steps, f"problem {p}: status {st[p]}, steps {n[p]} / {no}"
        if prec == 0:
            assert t[p] == to, f"problem {p}"
            assert max_rel(pb, so, sg) <= 1e-12, f"problem {p}"
            assert exact_fraction(pb, so, sg) == 1.0, f"problem {p} ({pb.name}) not bit-exact"
        else:
            from tenjin_amd.configs import RHO

            c = pb.layout.comp(RHO)
            assert abs(t[p] - to) <= 1e-5 * to, f"problem {p}"
            err = _rel_l1(sg[c], so[c])
            assert err <= 1e-5, f"problem {p} ({pb.name}): rho L1 {err:.3e}"
    b.close()


def test_full_ensemble_parity_and_properties():
    from tenjin_amd import GpuBatch, configs
    from tenjin_amd.configs import ALPHARHO

    probs = configs.ensemble()
    cells = sum(p.ncells for p in probs)
    print(f"\n[case] full ensemble: {len(probs)} problems, {cells} cells", flush=True)
    one = [q for q in range(len(probs)) if probs[q].dim == 1]
    two = [q for q in range(len(probs)) if probs[q].dim == 2]
    pick = sorted({max(one, key=lambda q: probs[q].ncells), min(one, key=lambda q: probs[q].ncells),
                   max(two, key=lambda q: probs[q].ncells), min(two, key=lambda q: probs[q].ncells)})
    # FP64, 20 steps: the picked problems against the oracle
    b = GpuBatch(probs)
    b.init_from_prims()
    b.begin_run()
    b.advance(20, 0)
    t, n, st = b.clocks()
    assert all(s == 0 for s in st) and all(k == 20 for k in n)
    for q in pick:
        pb = probs[q]
        print(f"[case] full ensemble problem {q} {pb.name} n={pb.n[:pb.dim]} vs oracle", flush=True)
        so, to, no = _oracle_run(pb, 20)
        sg = b.state(q)
        assert (t[q], n[q]) == (to, no)
        assert exact_fraction(pb, so, sg) == 1.0, f"problem {q} ({pb.name}) not bit-exact"
    b.close()
    for prec in (0, 1):
        print(f"[case] full ensemble: 200 steps prec={prec}, mass per material", flush=True)
        b = GpuBatch(probs)
        b.init_from_prims()
        b.begin_run()
        b.advance(200, prec)
        t, n, st = b.clocks()
        assert all(s == 0 for s in st) and all(k == 200 for k in n), "a problem stopped early"
        checked = 0
        for q, pb in enumerate(probs):
            s = b.state(q)
            assert np.all(np.isfinite(s[pb.layout.state_comps()])), f"problem {q}"
            if not _closed(pb, s):
                continue
            L = pb.layout
            for m in range(2):
                m0 = pb.prims[L.comp(0, m)] * pb.prims[L.comp(2, m)]  # alpha_m rho_m of the initial state
                got = s[L.comp(ALPHARHO, m)].sum()
                assert abs(got - m0.sum()) <= 1e-12 * m0.sum(), f"prec {prec} problem {q} material {m}: {got - m0.sum():.3e}"
            checked += 1
        assert checked >= len(probs) // 2
        b.close()


def _closed(pb, s):
    """No mass has crossed the boundary: the normal velocity of every boundary
    cell is exactly 0 (x ends; the y walls of a 2D problem)."""
    from tenjin_amd.configs import VELOCITY

    L = pb.layout
    ux = s[L.comp(VELOCITY, 0, 0)]
    if not (np.all(ux[..., 0] == 0.0) and np.all(ux[..., -1] == 0.0)):
        return False
    if pb.dim == 2:
        uy = s[L.comp(VELOCITY, 0, 1)]
        return bool(np.all(uy[:, 0, :] == 0.0) and np.all(uy[:, -1, :] == 0.0))
    return True
