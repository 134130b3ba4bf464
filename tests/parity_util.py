"""Comparison helpers for the GPU-vs-oracle parity tests."""
from __future__ import annotations

import numpy as np


def field_errors(pb, ref: np.ndarray, got: np.ndarray, comps=None):
    """Per-component max-norm error relative to the component's max-norm,
    over the persistent state components (conserved, primitive and rho)."""
    L = pb.layout
    comps = comps if comps is not None else L.state_comps()
    out = {}
    for c in comps:
        r = ref[c]
        g = got[c]
        scale = np.max(np.abs(r))
        err = np.max(np.abs(g - r))
        out[L.names[c]] = (err / scale) if scale > 0 else err
    return out


def max_rel(pb, ref, got, comps=None) -> float:
    e = field_errors(pb, ref, got, comps)
    return max(e.values()) if e else 0.0


def exact_fraction(pb, ref, got, comps=None) -> float:
    L = pb.layout
    comps = comps if comps is not None else L.state_comps()
    eq = [np.mean(ref[c] == got[c]) for c in comps]
    return float(np.mean(eq))
