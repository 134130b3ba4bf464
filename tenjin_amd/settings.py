"""Reference-format settings decks (libconfig syntax, the keys read by
initialise.cpp:125-385) -> a Problem, with a pluggable initial-condition
geometry in place of the reference's hard-coded chooseStateBasedOnInitialCondition
(initialise.cpp:387-581).

Supported: fluid materials with EOS "MieGruneisen" (stiffened gas /
constant-Gamma Mie-Gruneisen: adiabaticIndex, pref, eref, CV), one of them
optionally a mixture (mixture = 1: MixtureEOS with the products'
adiabaticIndexmix, prefmix, erefmix, CVmix, initialise.cpp:183-193, and per
state lambda and rhoa, :245-249); solid materials with EOS "UsUpSolid" (rho0,
Gamma0, c0, s, G0, yieldStress; this project's Us-up Mie-Gruneisen solid,
which has no reference class); solid materials with EOS "RomenskiiSolid"
(adiabaticIndex, pref, eref, CV, rho0, K0, EOSalpha, EOSbeta, G0,
yieldStress) and "WilkinsSolid" (adiabaticIndex, pref, eref, CV, rho0,
e1 .. e5, G0, yieldStress), the keys of initialise.cpp:149-172, 205-215;
RADIAL and REACTIVE (initialise.cpp:332-335) switch the geometric and the
reactive source on.  A second mixture is rejected with an error.  Host-side
set-up only.
"""
from __future__ import annotations

import math
import re
from typing import Callable, Optional

import numpy as np

from .configs import (ALPHA, EPSILON, LAMBDA, P, RHO_K, RHO_MIX, V_TENSOR, VELOCITY, Material, Problem,
                      _empty_prims)

# ---------------------------------------------------------------- libconfig subset
_TOKEN = re.compile(r"""
    (?P<ws>\s+|\#[^\n]*|//[^\n]*|/\*.*?\*/)
  | (?P<str>"(?:[^"\\]|\\.)*")
  | (?P<num>[-+]?(?:0[xX][0-9a-fA-F]+|(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?)L?)
  | (?P<bool>\b(?:true|false|TRUE|FALSE|True|False)\b)
  | (?P<name>[A-Za-z\*][-A-Za-z0-9_\*]*)
  | (?P<punct>[=:;,\{\}\[\]\(\)])
""", re.VERBOSE | re.DOTALL)


def _tokens(text: str):
    pos = 0
    out = []
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m:
            raise ValueError(f"settings: cannot parse near {text[pos:pos + 30]!r}")
        pos = m.end()
        kind = m.lastgroup
        if kind == "ws":
            continue
        out.append((kind, m.group(kind)))
    return out


class _Parser:
    def __init__(self, toks):
        self.t = toks
        self.i = 0

    def peek(self):
        return self.t[self.i] if self.i < len(self.t) else (None, None)

    def take(self, val=None):
        tok = self.peek()
        if val is not None and tok[1] != val:
            raise ValueError(f"settings: expected {val!r}, found {tok[1]!r}")
        self.i += 1
        return tok

    def settings(self, end=None):
        out = {}
        while self.peek()[1] != end and self.peek()[0] is not None:
            kind, name = self.take()
            if kind != "name":
                raise ValueError(f"settings: expected a setting name, found {name!r}")
            if self.peek()[1] not in ("=", ":"):
                raise ValueError(f"settings: expected '=' after {name}")
            self.take()
            out[name] = self.value()
            if self.peek()[1] in (";", ","):
                self.take()
        return out

    def value(self):
        kind, tok = self.peek()
        if tok == "{":
            self.take()
            v = self.settings("}")
            self.take("}")
            return v
        if tok in ("[", "("):
            close = "]" if tok == "[" else ")"
            self.take()
            items = []
            while self.peek()[1] != close:
                items.append(self.value())
                if self.peek()[1] == ",":
                    self.take()
            self.take(close)
            return items
        self.take()
        if kind == "str":
            s = bytes(tok[1:-1], "utf-8").decode("unicode_escape")
            while self.peek()[0] == "str":  # adjacent strings concatenate
                s += bytes(self.take()[1][1:-1], "utf-8").decode("unicode_escape")
            return s
        if kind == "bool":
            return tok.lower() == "true"
        if kind == "num":
            t = tok.rstrip("L")
            if t.lower().startswith(("0x", "+0x", "-0x")):
                return int(t, 16)
            return float(t) if any(c in t for c in ".eE") else int(t)
        raise ValueError(f"settings: unexpected {tok!r}")


def parse(text: str) -> dict:
    """Parses libconfig text (settings, groups, arrays, lists, strings,
    numbers, booleans, comments) into nested dicts and lists."""
    return _Parser(_tokens(text)).settings()


# ---------------------------------------------------------------- geometry
def riemann_1d(i, j, k, pb, cfg):
    """The 1D Riemann-problem rule of initialise.cpp:393-402: state 0 when
    i < (int)((interface / xdomainLength) * numberOfxCells), else state 1
    (applied along x in any dimension)."""
    return np.where(i < int((cfg["interface"] / pb.L[0]) * pb.n[0]), 0, 1)


GEOMETRIES = {"riemann1d": riemann_1d}


def _stretch(F):
    """V = sqrt(F F^T) / det(.)^(1/3) (getStretchTensor and the normalisation
    of initialise.cpp:31-46), by an eigen-decomposition of F F^T."""
    F = np.asarray(F, dtype=np.float64).reshape(3, 3)
    w, q = np.linalg.eigh(F @ F.T)
    V = (q * np.sqrt(w)) @ q.T
    return V * np.linalg.det(V) ** (-1.0 / 3.0)


def problem_from_settings(cfg, max_grid_size: int = 64, dim: Optional[int] = None,
                          geometry: "str | Callable" = "riemann1d", name: str = "settings") -> Problem:
    """Problem from a parsed deck (or deck text).  geometry: a name in
    GEOMETRIES or a callable (i, j, k, pb, cfg) -> state index per cell
    (numpy index arrays of the level's shape)."""
    if isinstance(cfg, str):
        cfg = parse(cfg)
    n = [int(cfg.get("numberOfxCells", 1)), int(cfg.get("numberOfyCells", 1)), int(cfg.get("numberOfzCells", 1))]
    L = [float(cfg.get("xdomainLength", 1.0)), float(cfg.get("ydomainLength", 1.0)), float(cfg.get("zdomainLength", 1.0))]
    if dim is None:
        dim = 3 if n[2] > 1 else (2 if n[1] > 1 else 1)
    nm = int(cfg["materials"])
    if int(cfg.get("mixtures", 0)) > 1:
        raise ValueError("settings: at most one mixture (every mixture shares mixtureIndex 0, initialise.cpp:187)")
    solid_run = int(cfg.get("SOLID", 0))
    plastic_run = int(cfg.get("PLASTIC", 0))
    mats = []
    for m in range(nm):
        md = cfg["listOfMaterials"][m]
        eos = md.get("EOS", "")
        phase = md.get("material", "fluid")
        if int(md.get("mixture", 0)):
            if eos != "MieGruneisen" or phase != "fluid":
                raise ValueError(f"settings: material {m}: only a MieGruneisen fluid can be a mixture")
            mats.append(Material.mixture_mg(
                (float(md["adiabaticIndex"]), float(md["pref"]), float(md["eref"]), float(md["CV"])),
                (float(md["adiabaticIndexmix"]), float(md["prefmix"]), float(md["erefmix"]), float(md["CVmix"]))))
        elif eos == "MieGruneisen" and phase == "fluid":
            mats.append(Material(gamma=float(md["adiabaticIndex"]), pref=float(md["pref"]), eref=float(md["eref"]),
                                 CV=float(md["CV"])))
        elif eos == "UsUpSolid" and phase == "solid":
            mats.append(Material.usup_solid(float(md["rho0"]), float(md["Gamma0"]), float(md["c0"]), float(md["s"]),
                                            float(md["G0"]), float(md.get("yieldStress", 0.0)), bool(plastic_run)))
        elif eos in ("RomenskiiSolid", "WilkinsSolid"):
            # getMaterialParameters (initialise.cpp:125-235): a solid EOS cannot be a
            # mixture; G0 (and yieldStress on PLASTIC runs) come with phase "solid"
            if phase != "solid":
                raise ValueError(f"settings: material {m}: {eos} needs material = \"solid\"")
            base = (float(md["adiabaticIndex"]), float(md["pref"]), float(md["eref"]), float(md["CV"]))
            Y = float(md["yieldStress"]) if plastic_run else 0.0
            if eos == "RomenskiiSolid":
                mats.append(Material.romenskii_solid(*base, float(md["rho0"]), float(md["K0"]), float(md["EOSalpha"]),
                                                     float(md["EOSbeta"]), float(md["G0"]), Y, bool(plastic_run)))
            else:
                mats.append(Material.wilkins_solid(*base, float(md["rho0"]),
                                                   tuple(float(md[f"e{q}"]) for q in range(1, 6)), float(md["G0"]), Y,
                                                   bool(plastic_run)))
        else:
            raise ValueError(f"settings: material {m}: EOS {eos!r} for a {phase} is not part of this path")
    low = [int(v) for v in cfg.get("lowBoundary", [0, 0, 0])] + [0, 0, 0]
    high = [int(v) for v in cfg.get("highBoundary", [0, 0, 0])] + [0, 0, 0]
    nn = tuple(n[a] if a < dim else 1 for a in range(3))
    pb = Problem(name, dim, nn, tuple(L), max_grid_size, mats, SOLID=solid_run, THINC=int(cfg.get("THINC", 0)),
                 PLASTIC=plastic_run, MUSCL=int(cfg.get("MUSCL", 1)), THINCbeta=float(cfg.get("THINCbeta", 1.6)),
                 CFL=float(cfg.get("CFL", 0.5)), bc_lo=tuple(low[:3]), bc_hi=tuple(high[:3]),
                 final_time=float(cfg["finalTime"]) if "finalTime" in cfg else None, shift_pressure=1,
                 RADIAL=int(cfg.get("RADIAL", 0)), REACTIVE=int(cfg.get("REACTIVE", 0)))
    if sum(mt.mixture for mt in mats) > 1:
        raise ValueError("settings: at most one mixture material")
    # states (initialise.cpp:236-263) and the per-cell choice
    states = cfg["listOfStates"][: int(cfg.get("states", len(cfg["listOfStates"])))]
    nz, ny, nx = pb.shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    geo = GEOMETRIES[geometry] if isinstance(geometry, str) else geometry
    sel = np.asarray(geo(i, j, k, pb, cfg))
    U = _empty_prims(pb)
    Lay = pb.layout
    for s, st in enumerate(states):
        mask = sel == s
        if not mask.any():
            continue
        # getState (initialise.cpp:237-263) keeps one F per state, and each
        # solid material's entry overwrites it: the last solid's F sets V and
        # the density of every solid (initialise.cpp:27-46, 74-77)
        Fs = None
        for m in range(nm):
            if mats[m].solid:
                Fs = np.asarray(st["material"][m]["F"], dtype=np.float64).reshape(3, 3)
        for m in range(nm):
            sm = st["material"][m]
            U[Lay.comp(ALPHA, m)][mask] = float(sm["alpha"])
            if mats[m].solid:
                U[Lay.comp(RHO_K, m)][mask] = mats[m].rho0 / np.linalg.det(Fs)  # setRhoFromDeformationTensor
            elif mats[m].mixture:  # initialise.cpp:66-71: rho_k and both phase densities = rhoa
                U[Lay.comp(RHO_K, m)][mask] = float(sm["rhoa"])
                U[Lay.comp(RHO_MIX, m, 0)][mask] = float(sm["rhoa"])
                U[Lay.comp(RHO_MIX, m, 1)][mask] = float(sm["rhoa"])
                U[Lay.comp(LAMBDA, m)][mask] = float(sm["lambda"])
            else:
                U[Lay.comp(RHO_K, m)][mask] = float(sm["rho"])
            if plastic_run and solid_run:
                U[Lay.comp(EPSILON, m)][mask] = 0.0
        U[Lay.comp(P)][mask] = float(st["p"])
        for r, key in enumerate(("u", "v", "w")):
            U[Lay.comp(VELOCITY, 0, r)][mask] = float(st.get(key, 0.0))
        if solid_run:
            V = _stretch(Fs) if Fs is not None else np.eye(3)
            for r in range(3):
                for c in range(3):
                    U[Lay.comp(V_TENSOR, 0, r, c)][mask] = V[r, c]
    pb.prims = U
    return pb
