"""Workload definitions and initial conditions for the per-cell update.

A ``Problem`` carries what the reference reads from its settings file
(initialise.cpp:310-375: domain, CFL, flags, materials, boundaries) plus the
initial primitive field.  Initial conditions follow the reference's
``initial_conditions`` (initialise.cpp:20-109): per cell, volume fractions,
phase densities, velocity, pressure (and V for solids) are set, optionally the
thermal pressure is shifted to total pressure (initialise.cpp:104), then
``primitiveToConservative`` fills the conserved fields.

Stiffened gas maps onto the reference's MieGruneisenEOS as
``adiabaticIndex = gamma, pref = -gamma * p_inf, eref = 0``
(equationofstate.cpp:37-64), so a physical pressure is set directly.

Random draws are 53-bit uniforms from the 32-bit Mersenne twister seeded with
init_genrand(seed) (numpy's legacy RandomState): u = ((a >> 5) 2^26 + (b >> 6)) 2^-53
from two consecutive outputs a, b, which C++ ``std::mt19937(seed)`` reproduces.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

# Variable enum of structdefinitions.h:10-32
ALPHA, ALPHARHO, RHO_K, RHO, RHOU, TOTAL_E, VELOCITY, P, SOUNDSPEED, USTAR = range(10)
RHO_MIX, ALPHARHOLAMBDA, LAMBDA, SIGMA, V_TENSOR, VSTAR, DEVH, HJ2, EPSILON, ALPHARHOEPSILON = range(10, 20)

ABSENT_ALPHA = 1e-6


class Layout:
    """Component map of accesspattern.cpp:44-79 (numberOfMixtures 0 or 1)."""

    def __init__(self, nmat: int, solid: bool, plastic: bool, nmix: int = 0):
        self.start = {}
        self.names: List[str] = []
        self.cons: List[tuple] = []
        self.prim: List[tuple] = []
        n = 0

        def add(base, kind, var, nm, nr, nc):
            nonlocal n
            self.start[var] = n
            n += nm * nr * nc
            for m in range(nm):
                for r in range(nr):
                    for c in range(nc):
                        self.names.append(f"{base}{m}_{r}{c}")
                        if kind in ("cons", "both"):
                            self.cons.append((var, m, r, c))
                        if kind in ("prim", "both"):
                            self.prim.append((var, m, r, c))

        add("alpha", "both", ALPHA, nmat, 1, 1)
        add("alphaRho", "cons", ALPHARHO, nmat, 1, 1)
        add("rho_k", "prim", RHO_K, nmat, 1, 1)
        add("rho", "none", RHO, 1, 1, 1)
        add("rhoU", "cons", RHOU, 1, 3, 1)
        add("u", "prim", VELOCITY, 1, 3, 1)
        add("E", "cons", TOTAL_E, 1, 1, 1)
        add("p", "prim", P, 1, 1, 1)
        add("a", "none", SOUNDSPEED, 1, 1, 1)
        add("uStar", "none", USTAR, 1, 1, 1)
        add("sigma", "none", SIGMA, 1, 3, 3)
        if nmix:
            add("rhomix", "none", RHO_MIX, nmix, 2, 1)
            add("lambda", "prim", LAMBDA, nmix, 1, 1)
            add("alpharholambda", "cons", ALPHARHOLAMBDA, nmix, 1, 1)
        self.nmix = nmix
        if solid:
            add("V", "both", V_TENSOR, 1, 3, 3)
            add("VStar", "none", VSTAR, 1, 3, 3)
            add("devH", "none", DEVH, 1, 3, 3)
            add("HenckyJ2", "none", HJ2, 1, 1, 1)
            if plastic:
                add("epsilon", "prim", EPSILON, nmat, 1, 1)
                add("alphaRhoEpsilon", "cons", ALPHARHOEPSILON, nmat, 1, 1)
        self.ncomp = n

    def comp(self, var, m=0, r=0, c=0) -> int:
        if var in (ALPHA, ALPHARHO, RHO_K, EPSILON, ALPHARHOEPSILON):
            return self.start[var] + m
        if var in (RHOU, VELOCITY):
            return self.start[var] + r
        if var in (SIGMA, V_TENSOR, VSTAR, DEVH):
            return self.start[var] + 3 * r + c
        if var == RHO_MIX:  # every mixture uses mixtureIndex 0 (initialise.cpp:187)
            return self.start[var] + r
        return self.start[var]

    def cons_comps(self) -> List[int]:
        return [self.comp(*s) for s in self.cons]

    def prim_comps(self) -> List[int]:
        return [self.comp(*s) for s in self.prim]

    def state_comps(self) -> List[int]:
        """Persistent per-cell state: conserved and primitive fields plus rho
        (and the phase densities of a mixture, which the next sound speed reads)."""
        extra = {self.start[RHO]}
        if self.nmix:
            extra |= {self.start[RHO_MIX], self.start[RHO_MIX] + 1}
        return sorted(set(self.cons_comps()) | set(self.prim_comps()) | extra)


@dataclasses.dataclass
class Material:
    """One material: kind 0 = stiffened gas / MieGruneisenEOS, kind 1 = Us-up Mie-Gruneisen solid."""

    gamma: float = 1.4
    pref: float = 0.0
    eref: float = 0.0
    CV: float = 1.0
    eos: int = 0
    solid: int = 0
    plastic: int = 0
    rho0: float = 1.0
    c0: float = 0.0
    s: float = 0.0
    G0: float = 0.0
    yield_stress: float = 0.0
    # MixtureEOS (equationofstate.cpp:116-376): the fields above are the
    # reactants' MieGruneisen EOS ("first"), these the products' ("second")
    mixture: int = 0
    gamma2: float = 1.4
    pref2: float = 0.0
    eref2: float = 0.0
    CV2: float = 1.0
    # RomenskiiSolidEOS (eos 2) and WilkinsSolidEOS (eos 3), with rho0 and G0
    K0: float = 0.0
    eos_alpha: float = 0.0
    eos_beta: float = 0.0
    e1: float = 0.0
    e2: float = 0.0
    e3: float = 0.0
    e4: float = 0.0
    e5: float = 0.0

    @staticmethod
    def stiffened(gamma: float, p_inf: float = 0.0) -> "Material":
        return Material(gamma=gamma, pref=-gamma * p_inf)

    @staticmethod
    def mixture_mg(a: tuple, b: tuple) -> "Material":
        """Reactant/product mixture; a, b = (adiabaticIndex, pref, eref, CV)."""
        return Material(gamma=a[0], pref=a[1], eref=a[2], CV=a[3], mixture=1,
                        gamma2=b[0], pref2=b[1], eref2=b[2], CV2=b[3])

    @staticmethod
    def romenskii_solid(gamma, pref, eref, CV, rho0, K0, alpha, beta, G, Y=0.0, plastic=True) -> "Material":
        """RomenskiiSolidEOS (equationofstate.cpp:403-495): adiabaticIndex,
        pref, eref, CV, rho0, K0, EOSalpha, EOSbeta, G0 as in a settings deck."""
        return Material(gamma=gamma, pref=pref, eref=eref, CV=CV, eos=2, solid=1, plastic=int(plastic), rho0=rho0,
                        G0=G, yield_stress=Y, K0=K0, eos_alpha=alpha, eos_beta=beta)

    @staticmethod
    def wilkins_solid(gamma, pref, eref, CV, rho0, e, G, Y=0.0, plastic=True) -> "Material":
        """WilkinsSolidEOS (equationofstate.cpp:502-596): adiabaticIndex, pref,
        eref, CV, rho0, e1 .. e5, G0 as in a settings deck."""
        return Material(gamma=gamma, pref=pref, eref=eref, CV=CV, eos=3, solid=1, plastic=int(plastic), rho0=rho0,
                        G0=G, yield_stress=Y, e1=e[0], e2=e[1], e3=e[2], e4=e[3], e5=e[4])

    @staticmethod
    def usup_solid(rho0, Gamma0, c0, s, G, Y, plastic=True) -> "Material":
        return Material(gamma=Gamma0 + 1.0, eos=1, solid=1, plastic=int(plastic), rho0=rho0,
                        c0=c0, s=s, G0=G, yield_stress=Y)


@dataclasses.dataclass
class Problem:
    name: str
    dim: int
    n: tuple
    L: tuple
    max_grid_size: int
    materials: List[Material]
    SOLID: int = 0
    THINC: int = 0
    PLASTIC: int = 0
    MUSCL: int = 1
    RADIAL: int = 0  # geometric (cylindrical / spherical) source, radial.cpp
    REACTIVE: int = 0  # pressure-based burn of the mixture materials, reactive.cpp
    THINCbeta: float = 1.6
    CFL: float = 0.5
    bc_lo: tuple = (0, 0, 0)
    bc_hi: tuple = (0, 0, 0)
    nghost: int = 2
    final_time: Optional[float] = None
    steps: Optional[int] = None
    shift_pressure: int = 0
    prims: Optional[np.ndarray] = None  # [ncomp][nz][ny][nx]

    @property
    def nmat(self) -> int:
        return len(self.materials)

    @property
    def layout(self) -> Layout:
        return Layout(self.nmat, bool(self.SOLID), bool(self.PLASTIC), self.nmix)

    @property
    def nmix(self) -> int:
        return 1 if any(m.mixture for m in self.materials) else 0

    @property
    def shape(self):
        nx = self.n[0]
        ny = self.n[1] if self.dim > 1 else 1
        nz = self.n[2] if self.dim > 2 else 1
        return nz, ny, nx

    @property
    def ncells(self) -> int:
        nz, ny, nx = self.shape
        return nx * ny * nz

    def dx(self):
        return [self.L[a] / self.n[a] if a < self.dim else 1.0 for a in range(3)]


def _empty_prims(pb: Problem) -> np.ndarray:
    return np.zeros((pb.layout.ncomp,) + pb.shape, dtype=np.float64)


def _set_two_phase(pb: Problem, U: np.ndarray, mask: np.ndarray, rho0: float, rho1: float,
                   u: np.ndarray, p: np.ndarray):
    """mask True -> material 0 present (alpha0 = 1 - 1e-6), else material 1."""
    L = pb.layout
    a0 = np.where(mask, 1.0 - ABSENT_ALPHA, ABSENT_ALPHA)
    U[L.comp(ALPHA, 0)] = a0
    U[L.comp(ALPHA, 1)] = 1.0 - a0
    U[L.comp(RHO_K, 0)] = rho0
    U[L.comp(RHO_K, 1)] = rho1
    for r in range(3):
        U[L.comp(VELOCITY, 0, r)] = u[r]
    U[L.comp(P)] = p


def shock_tube_1d(n: int = 1000, interface: float = 0.7, max_grid_size: int = 1 << 30) -> Problem:
    """Two-material stiffened-gas shock tube (water-like / air), x in [0, 1].

    State rule of initialise.cpp:393-402: cell i takes the left state when
    i < (int)((interface / L) * n).
    """
    mats = [Material.stiffened(4.4, 6e8), Material.stiffened(1.4, 0.0)]
    pb = Problem("shock_tube_1d", 1, (n, 1, 1), (1.0, 1.0, 1.0), max_grid_size, mats,
                 final_time=2.4e-4)
    U = _empty_prims(pb)
    i = np.arange(n)
    left = (i < int((interface / 1.0) * n))[None, None, :]
    zero = np.zeros((1, 1, n))
    p = np.where(left, 1e9, 1e5)
    _set_two_phase(pb, U, left, 1000.0, 50.0, [zero, zero, zero], p)
    pb.prims = U
    return pb


def rankine_hugoniot(gamma: float, rho1: float, p1: float, mach: float):
    """Post-shock state of a shock of Mach number `mach` moving into gas at rest."""
    c1 = math.sqrt(gamma * p1 / rho1)
    m2 = mach * mach
    p2 = p1 * (1.0 + 2.0 * gamma / (gamma + 1.0) * (m2 - 1.0))
    rho2 = rho1 * (gamma + 1.0) * m2 / ((gamma - 1.0) * m2 + 2.0)
    w = mach * c1
    u2 = w * (1.0 - rho1 / rho2)
    return rho2, u2, p2


def _mt_uniform(seed: int, count: int) -> np.ndarray:
    return np.random.RandomState(seed).random_sample(count)


def shock_inclusion_2d(nx: int = 2048, ny: int = 1024, box: int = 64, seed: int = 1,
                       steps: int = 500, thinc: int = 1, shock_x: float = 0.5,
                       drift=(0.0, 0.0)) -> Problem:
    """Mach 1.22 shock in gas (gamma 1.4) hitting a light circular inclusion
    (gamma 1.67, rho 0.138) whose radius carries 8 seeded Fourier modes.

    Domain [0, 2] x [0, 1] (dx = dy = 1/ny for the full-size config); x
    transmissive, y reflective.  Cell centres decide the material.  `drift`
    adds a uniform velocity to every cell (test variants that move the
    interface from the first step).
    """
    mats = [Material.stiffened(1.4, 0.0), Material.stiffened(1.67, 0.0)]
    Lx, Ly = 2.0, 1.0
    pb = Problem("shock_inclusion_2d", 2, (nx, ny, 1), (Lx, Ly, 1.0), box, mats,
                 THINC=thinc, THINCbeta=1.6, bc_lo=(0, 1, 0), bc_hi=(0, 1, 0), steps=steps)
    U = _empty_prims(pb)
    x = (np.arange(nx) + 0.5) * (Lx / nx)
    y = (np.arange(ny) + 0.5) * (Ly / ny)
    X, Y = np.meshgrid(x, y)  # [ny][nx]
    uu = _mt_uniform(seed, 16)
    amp = -0.01 + 0.02 * uu[:8]
    phase = 2.0 * math.pi * uu[8:]
    xc, yc, r0 = 0.8, 0.5, 0.2 * Ly
    th = np.arctan2(Y - yc, X - xc)
    rr = r0 * (1.0 + sum(amp[k] * np.cos((k + 1) * th + phase[k]) for k in range(8)))
    inside = np.hypot(X - xc, Y - yc) < rr
    rho2, u2, p2 = rankine_hugoniot(1.4, 1.0, 1e5, 1.22)
    shocked = X < shock_x
    gas_rho = np.where(shocked, rho2, 1.0)
    ux = np.where(shocked, u2, 0.0) + drift[0]
    uy = np.zeros_like(X) + drift[1]
    p = np.where(shocked, p2, 1e5)
    mask = ~inside  # material 0 (gas) present outside the inclusion
    zero = np.zeros_like(X)
    _set_two_phase(pb, U, mask[None], gas_rho[None], 0.138, [ux[None], uy[None], zero[None]], p[None])
    pb.prims = U
    return pb


def shock_inclusion_3d(n=(32, 24, 20), box: int = 8, drift=(80.0, 40.0, 20.0),
                       bc_lo=(0, 1, 0), bc_hi=(0, 0, 1)) -> Problem:
    """3D counterpart of the shock/inclusion case for the fluid path (no THINC,
    which the reference defines in 1D/2D only): a Mach 1.22 shock in gas
    approaching a light spherical inclusion, with a uniform drift."""
    mats = [Material.stiffened(1.4, 0.0), Material.stiffened(1.67, 0.0)]
    Lx = 1.0
    h = Lx / n[0]
    pb = Problem("shock_inclusion_3d", 3, tuple(n), (Lx, h * n[1], h * n[2]), box, mats,
                 bc_lo=bc_lo, bc_hi=bc_hi)
    U = _empty_prims(pb)
    nz, ny, nx = pb.shape
    x = (np.arange(nx) + 0.5) * h
    y = (np.arange(ny) + 0.5) * h
    z = (np.arange(nz) + 0.5) * h
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    inside = (X - 0.55) ** 2 + (Y - 0.5 * h * ny) ** 2 + (Z - 0.5 * h * nz) ** 2 < (0.25 * h * ny) ** 2
    rho2, u2, p2 = rankine_hugoniot(1.4, 1.0, 1e5, 1.22)
    shocked = X < 0.3
    gas_rho = np.where(shocked, rho2, 1.0)
    ux = np.where(shocked, u2, 0.0) + drift[0]
    uy = np.zeros_like(X) + drift[1]
    uz = np.zeros_like(X) + drift[2]
    p = np.where(shocked, p2, 1e5)
    _set_two_phase(pb, U, ~inside, gas_rho, 0.138, [ux, uy, uz], p)
    pb.prims = U
    return pb


COPPER = dict(rho0=8930.0, Gamma0=2.0, c0=3940.0, s=1.49, G=45e9, Y=90e6)
ALUMINIUM = dict(rho0=2703.0, Gamma0=2.0, c0=5328.0, s=1.338, G=26e9, Y=300e6)


def plate_impact(dim: int = 3, n: int = 256, box: int = 64, seed: int = 3, speed: float = 400.0,
                 steps: int = 300, length: float = 25.6e-3, tilt_deg: float = 5.0,
                 plastic: bool = True, cu=None, al=None) -> Problem:
    """Copper flyer on aluminium across a plane interface through the domain
    centre tilted about z by `tilt_deg` from the x normal; Us-up
    Mie-Gruneisen solids with perfect plasticity; V = I, thermal p = 0,
    +-0.1 % seeded density noise.  All boundaries transmissive."""
    cu = dict(COPPER, **(cu or {}))
    al = dict(ALUMINIUM, **(al or {}))
    mats = [Material.usup_solid(cu["rho0"], cu["Gamma0"], cu["c0"], cu["s"], cu["G"], cu["Y"], plastic),
            Material.usup_solid(al["rho0"], al["Gamma0"], al["c0"], al["s"], al["G"], al["Y"], plastic)]
    nn = (n, n, n if dim > 2 else 1)
    pb = Problem(f"plate_impact_{dim}d", dim, nn, (length, length, length), box, mats,
                 SOLID=1, PLASTIC=int(plastic), steps=steps, shift_pressure=1)
    L = pb.layout
    U = _empty_prims(pb)
    nz, ny, nx = pb.shape
    h = length / n
    x = (np.arange(nx) + 0.5) * h
    y = (np.arange(ny) + 0.5) * h
    z = (np.arange(nz) + 0.5) * h
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    t = math.radians(tilt_deg)
    cu_side = (X - 0.5 * length) * math.cos(t) + (Y - 0.5 * length) * math.sin(t) < 0.0
    noise = _mt_uniform(seed, pb.ncells).reshape(pb.shape)
    fac = 1.0 + 1e-3 * (2.0 * noise - 1.0)
    a0 = np.where(cu_side, 1.0 - ABSENT_ALPHA, ABSENT_ALPHA)
    U[L.comp(ALPHA, 0)] = a0
    U[L.comp(ALPHA, 1)] = 1.0 - a0
    U[L.comp(RHO_K, 0)] = cu["rho0"] * fac
    U[L.comp(RHO_K, 1)] = al["rho0"] * fac
    U[L.comp(VELOCITY, 0, 0)] = np.where(cu_side, speed, 0.0)
    U[L.comp(P)] = 0.0
    for r in range(3):
        U[L.comp(V_TENSOR, 0, r, r)] = 1.0
    pb.prims = U
    return pb


# The plate impact's copper and aluminium as the reference's other solid EOSs:
# RomenskiiSolidEOS copper (K0 = rho0 c0^2, EOSalpha 1.3, EOSbeta 2.2) and a
# WilkinsSolidEOS aluminium whose polynomial has e_c = p_c = 0 at rho0 and
# bulk modulus rho0 c0^2 (e1 + e2 + e3 + e4 = 0, 2 e2 + e3 - e4 - e5 = 0,
# 6 e2 + 2 e3 - e5 = K).
def romenskii_copper(plastic: bool = True) -> Material:
    cu = COPPER
    return Material.romenskii_solid(cu["Gamma0"] + 1.0, 0.0, 0.0, 390.0, cu["rho0"], cu["rho0"] * cu["c0"] ** 2,
                                    1.3, 2.2, cu["G"], cu["Y"], plastic)


def wilkins_aluminium(plastic: bool = True) -> Material:
    al = ALUMINIUM
    K = al["rho0"] * al["c0"] ** 2
    e = (0.5 * K, 0.5 * K, -0.85 * K, -0.15 * K, 0.3 * K)
    return Material.wilkins_solid(al["Gamma0"] + 1.0, 0.0, 0.0, 900.0, al["rho0"], e, al["G"], al["Y"], plastic)


def plate_impact_rw(dim: int = 2, n: int = 48, box: int = 24, **kw) -> Problem:
    """plate_impact with the Romenskii copper flyer and the Wilkins aluminium
    target (the densities, speeds and geometry unchanged)."""
    pb = plate_impact(dim=dim, n=n, box=box, **kw)
    pb.materials = [romenskii_copper(bool(pb.PLASTIC)), wilkins_aluminium(bool(pb.PLASTIC))]
    pb.name = f"plate_impact_rw_{dim}d"
    return pb


PLATE_PARAMS = ("Gamma0", "s", "Y")  # per material: Us-up Gruneisen coefficient, Hugoniot slope, yield stress


def plate_params(cu=None, al=None) -> list:
    """[Gamma0, s, Y] of copper then aluminium (the differentiated parameters)."""
    cu = dict(COPPER, **(cu or {}))
    al = dict(ALUMINIUM, **(al or {}))
    return [cu[k] for k in PLATE_PARAMS] + [al[k] for k in PLATE_PARAMS]


def plate_impact_differentiated(n: int = 512, box: int = 64, seed: int = 4, theta=None, steps: int = 200) -> Problem:
    """The 2D elastoplastic plate impact of the differentiated form: the
    copper/aluminium geometry of plate_impact in 2D, n x n cells, impact speed
    U(200, 600) from `seed` (the density noise uses the same seed).  theta:
    [Gamma0, s, Y] of copper then aluminium (default the tabulated values).
    The probe is u_x of the cell two in from the aluminium's far (x high)
    face at mid-height, recorded after every step."""
    speed = 200.0 + 400.0 * float(_mt_uniform(seed, 1)[0])
    th = plate_params() if theta is None else [float(v) for v in theta]
    cu = dict(zip(PLATE_PARAMS, th[:3]))
    al = dict(zip(PLATE_PARAMS, th[3:6]))
    pb = plate_impact(dim=2, n=n, box=box, seed=seed, speed=speed, steps=steps, cu=cu, al=al)
    pb.name = "plate_impact_differentiated_2d"
    pb.speed = speed
    return pb


def plate_probe(pb: Problem):
    """(component, i, j, k) of the differentiated plate impact's probe."""
    nz, ny, nx = pb.shape
    return pb.layout.comp(VELOCITY, 0, 0), nx - 3, ny // 2, 0


def plate_target_theta(seed: int = 5) -> list:
    """The parameters of the target trace: each tabulated value times
    1 + 0.05 U(-1, 1) from `seed`."""
    u = _mt_uniform(seed, 6)
    return [v * (1.0 + 0.05 * (2.0 * w - 1.0)) for v, w in zip(plate_params(), u)]


def underwater_explosion(dim: int = 1, n: int = 800, box: int = 1 << 30, radius: float = 0.2,
                         thinc: int = 0) -> Problem:
    """High-pressure gas bubble (gamma 1.4, rho 1270, p 8.29e8) in water
    (stiffened gas 4.4 / 6e8, rho 1000, p 1e5) about the origin of a radial
    coordinate, with the geometric source on (RADIAL):

    * dim 1: a sphere, x = r in [0, 1] (multiplier -2 in radial.cpp:11);
    * dim 2: (x, y) = (r, z) in [0, 1] x [0, 1], the bubble centred on the
      axis at z = 1/2 (cylindrical, multiplier -1).

    Reflective at r = 0 (x low), transmissive elsewhere.  Cell centres decide
    the material."""
    mats = [Material.stiffened(4.4, 6e8), Material.stiffened(1.4, 0.0)]
    nn = (n, n if dim > 1 else 1, 1)
    pb = Problem(f"underwater_explosion_{dim}d", dim, nn, (1.0, 1.0, 1.0), box, mats, THINC=thinc,
                 bc_lo=(1, 0, 0), bc_hi=(0, 0, 0), final_time=2e-4, RADIAL=1)
    U = _empty_prims(pb)
    nz, ny, nx = pb.shape
    r = (np.arange(nx) + 0.5) / nx
    z = (np.arange(ny) + 0.5) / ny if dim > 1 else np.zeros(1)
    Z, R = np.meshgrid(z, r, indexing="ij")  # [ny][nx]
    inside = (np.hypot(R, Z - 0.5) if dim > 1 else R) < radius
    water = ~inside[None]
    zero = np.zeros((1, ny, nx))
    _set_two_phase(pb, U, water, 1000.0, 1270.0, [zero, zero, zero], np.where(water, 1e5, 8.29e8))
    pb.prims = U
    return pb


def rod_impact_axisymmetric(n: int = 96, box: int = 32, speed: float = 300.0, radius: float = 0.3,
                            length: float = 10e-3, seed: int = 5) -> Problem:
    """(x, y) = (r, z) axisymmetric impact with the geometric source on: a
    copper rod (r < radius L, z > L/2) moving at -speed along z into
    aluminium at rest, which fills the rest of the domain; Us-up
    Mie-Gruneisen solids with perfect plasticity, V = I, thermal p = 0,
    +-0.1 % seeded density noise.  Reflective at r = 0, transmissive
    elsewhere."""
    cu, al = COPPER, ALUMINIUM
    mats = [Material.usup_solid(cu["rho0"], cu["Gamma0"], cu["c0"], cu["s"], cu["G"], cu["Y"], True),
            Material.usup_solid(al["rho0"], al["Gamma0"], al["c0"], al["s"], al["G"], al["Y"], True)]
    pb = Problem("rod_impact_axisymmetric", 2, (n, n, 1), (length, length, 1.0), box, mats,
                 SOLID=1, PLASTIC=1, RADIAL=1, bc_lo=(1, 0, 0), steps=60, shift_pressure=1)
    L = pb.layout
    U = _empty_prims(pb)
    h = length / n
    c = (np.arange(n) + 0.5) * h
    Z, R = np.meshgrid(c, c, indexing="ij")
    rod = (R < radius * length) & (Z > 0.5 * length)
    noise = _mt_uniform(seed, pb.ncells).reshape(pb.shape)
    fac = 1.0 + 1e-3 * (2.0 * noise - 1.0)
    a0 = np.where(rod, 1.0 - ABSENT_ALPHA, ABSENT_ALPHA)[None]
    U[L.comp(ALPHA, 0)] = a0
    U[L.comp(ALPHA, 1)] = 1.0 - a0
    U[L.comp(RHO_K, 0)] = cu["rho0"] * fac
    U[L.comp(RHO_K, 1)] = al["rho0"] * fac
    U[L.comp(VELOCITY, 0, 1)] = np.where(rod, -speed, 0.0)[None]
    U[L.comp(P)] = 0.0
    for q in range(3):
        U[L.comp(V_TENSOR, 0, q, q)] = 1.0
    pb.prims = U
    return pb


EXPLOSIVE_REACTANT = (4.4, -4.4 * 6e8, 0.0, 2000.0)  # adiabaticIndex, pref, eref, CV
EXPLOSIVE_PRODUCTS = (2.0, -2.0 * 2e8, 0.0, 1000.0)


def reactive_tube(dim: int = 1, n: int = 800, box: int = 1 << 30, thinc: int = 0, seed: int = 7) -> Problem:
    """A reactive mixture (stiffened-gas reactant and products,
    MixtureEOS) next to air, with the pressure-based burn of reactive.cpp:3-24
    (REACTIVE).  x in [0, 1]: an ignition region x < 0.1 of mostly burnt,
    high-pressure mixture (lambda 0.2, p 2e9), unburnt explosive (lambda
    seeded in [0.5, 1)) up to x = 0.6, air beyond.  In 2D the interface with
    the air is a seeded wavy line across y in [0, 1/2] (reflective in y).
    Cell centres decide the region; all x boundaries transmissive."""
    mats = [Material.mixture_mg(EXPLOSIVE_REACTANT, EXPLOSIVE_PRODUCTS), Material.stiffened(1.4, 0.0)]
    nn = (n, n // 2 if dim > 1 else 1, 1)
    pb = Problem(f"reactive_tube_{dim}d", dim, nn, (1.0, 0.5, 1.0), box, mats, THINC=thinc, REACTIVE=1,
                 bc_lo=(0, 1, 0), bc_hi=(0, 1, 0), final_time=1e-4)
    L = pb.layout
    U = _empty_prims(pb)
    nz, ny, nx = pb.shape
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) * (0.5 / ny)
    Y, X = np.meshgrid(y, x, indexing="ij")
    uu = _mt_uniform(seed, 8 + pb.ncells)
    edge = 0.6 + (0.02 * np.sin(2.0 * math.pi * (4.0 * Y + uu[0])) if dim > 1 else 0.0)
    explosive = (X < edge)[None]
    ignition = (X < 0.1)[None]
    lam = np.where(ignition, 0.2, 0.5 + 0.5 * uu[8:].reshape(pb.shape))
    zero = np.zeros((1, ny, nx))
    _set_two_phase(pb, U, explosive, 1000.0, 1.2, [zero, zero, zero],
                   np.where(ignition, 2e9, 1e5))
    U[L.comp(LAMBDA)] = lam
    # initialise.cpp:66-71: both phase densities start at the mixture's
    U[L.comp(RHO_MIX, 0, 0)] = U[L.comp(RHO_K, 0)]
    U[L.comp(RHO_MIX, 0, 1)] = U[L.comp(RHO_K, 0)]
    pb.prims = U
    return pb


def ensemble(seed: int = 6, n1d: int = 128, n2d: int = 128, n1d_range=(500, 8000), side_range=(64, 512),
             thinc: int = 1, steps: int = 200) -> List[Problem]:
    """Ensemble of independent two-material shock tubes: n1d 1D tubes with
    N ~ U{n1d_range} cells and n2d 2D square tubes with side ~ U{side_range}
    (planar interface at x = 0.5 with a one-mode sinusoidal perturbation in
    y).  Per problem: gamma ~ U[1.2, 5] and p_inf ~ U[0, 1e9] for each
    material, density ratio ~ logU[10, 1000] and pressure ratio
    ~ logU[10, 1e4] (the dense, high-pressure side on the left), unit domain,
    64-cell boxes, transmissive ends; 2D tubes reflective in y.  Even problems
    are stiffened-gas pairs; odd ones Mie-Gruneisen pairs (MieGruneisenEOS with
    a reference energy eref ~ U[-1e5, 1e5] J/kg per material as well).  Draws
    come from numpy's RandomState(seed) (MT19937)."""
    rs = np.random.RandomState(seed)
    out = []
    for q in range(n1d + n2d):
        dim = 1 if q < n1d else 2
        n = int(rs.randint(n1d_range[0], n1d_range[1] + 1)) if dim == 1 else int(
            rs.randint(side_range[0], side_range[1] + 1))
        g0, g1 = rs.uniform(1.2, 5.0, 2)
        pi0, pi1 = rs.uniform(0.0, 1e9, 2)
        drat = float(np.exp(rs.uniform(np.log(10.0), np.log(1000.0))))
        prat = float(np.exp(rs.uniform(np.log(10.0), np.log(1e4))))
        amp = float(rs.uniform(0.0, 0.05))
        er0, er1 = rs.uniform(-1e5, 1e5, 2)
        mats = [Material.stiffened(float(g0), float(pi0)), Material.stiffened(float(g1), float(pi1))]
        if q % 2:  # Mie-Gruneisen pair: the same closure with a reference energy
            mats[0].eref, mats[1].eref = float(er0), float(er1)
        nn = (n, n if dim == 2 else 1, 1)
        pb = Problem(f"ensemble_{q}_{dim}d", dim, nn, (1.0, 1.0, 1.0), 64, mats, THINC=thinc, steps=steps,
                     bc_lo=(0, 1, 0), bc_hi=(0, 1, 0))
        U = _empty_prims(pb)
        nz, ny, nx = pb.shape
        x = (np.arange(nx) + 0.5) / nx
        y = (np.arange(ny) + 0.5) / ny
        Y, X = np.meshgrid(y, x, indexing="ij")
        xi = 0.5 + (amp * np.sin(2.0 * np.pi * Y) if dim == 2 else 0.0)
        left = (X < xi)[None, :, :]
        zero = np.zeros(pb.shape)
        _set_two_phase(pb, U, left, drat * 1.0, 1.0, [zero, zero, zero],
                       np.where(left, 1e5 * prat, 1e5))
        pb.prims = U
        out.append(pb)
    return out
