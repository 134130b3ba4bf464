"""Writes the committed golden fixtures: small runs of the oracle, step by step.

The reference ships no golden vectors, tests or sample outputs, so these
fixtures are generated from the oracle (the CPU restatement) and pin it against
regressions; the GPU parity tests check the device against them as well.
Run:  python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.oracle import Oracle  # noqa: E402
from tenjin_amd import configs  # noqa: E402


def cases():
    yield "shock_tube_64", configs.shock_tube_1d(n=64), 12
    pb = configs.shock_inclusion_2d(nx=16, ny=16, box=8, thinc=1, shock_x=0.55, drift=(150.0, 60.0))
    pb.L = (0.25, 0.25, 1.0)
    pb.prims = configs.shock_inclusion_2d(nx=16, ny=16, box=8, thinc=1, shock_x=0.55, drift=(150.0, 60.0)).prims
    yield "inclusion_16x16_thinc", pb, 6


def run(name, pb, nsteps):
    o = Oracle(pb)
    comps = pb.layout.state_comps()
    states = [o.get()[comps]]
    dts = []
    for n in range(nsteps):
        dt = o.timestep()
        dts.append(dt)
        o.step(0.2 * dt if n < 5 else dt)
        states.append(o.get()[comps])
    return dict(comps=np.array(comps), states=np.array(states), dts=np.array(dts), prims=pb.prims)


if __name__ == "__main__":
    for name, pb, nsteps in cases():
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **run(name, pb, nsteps))
        print("wrote", name)
