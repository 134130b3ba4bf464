"""CPU tests of the oracle (the restatement the GPU is checked against).

The reference ships no tests or golden vectors, so each part of the oracle is
pinned by closed forms, invariants and independent computations, chosen so
that a dropped term, a sign or an index error fails one of them.
"""
from __future__ import annotations

import math
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from oracle.oracle import Oracle, lib
from tenjin_amd import configs as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def one_cell_problem(n=4, nmat=2, mats=None, dim=1):
    mats = mats or [C.Material.stiffened(4.4, 6e8), C.Material.stiffened(1.4, 0.0)]
    pb = C.Problem("unit", dim, (n, 1, 1), (1.0, 1.0, 1.0), 1 << 20, mats[:nmat])
    pb.prims = np.zeros((pb.layout.ncomp,) + pb.shape)
    return pb


# ---------------------------------------------------------------- math
@pytest.mark.parametrize("kind,lo,hi", [(0, -40.0, 40.0), (1, -4.0, 4.0), (2, -4.0, 4.0), (3, 0.05, 20.0),
                                         (4, 1e-6, 1e7), (5, 1e-3, 1e6), (6, 1e-4, 1e4)])
def test_project_math_within_ulps_of_libm(oracle_lib, kind, lo, hi):
    x = np.geomspace(lo, hi, 4001) if kind >= 4 else np.linspace(lo, hi, 4001)
    ref = {0: np.exp, 1: np.cosh, 2: np.tanh, 3: lambda v: v ** (-1.0 / 3.0), 4: np.log,
           5: lambda v: v ** 0.5, 6: lambda v: v ** 1.2}[kind](x)
    got = np.array([lib().tjo_math(kind, 0, float(v)) for v in x])
    libm = np.array([lib().tjo_math(kind, 1, float(v)) for v in x])
    nz = ref != 0
    # pow = exp(y log x) carries log's rounding times |y log x| (<= 11 here)
    ulps = 16 if kind == 6 else 4
    assert np.max(np.abs(got[nz] - ref[nz]) / np.abs(ref[nz])) < ulps * 2.0 ** -52
    assert np.max(np.abs(libm[nz] - ref[nz]) / np.abs(ref[nz])) < 4 * 2.0 ** -52


# ---------------------------------------------------------------- MUSCL
def test_vanleer_closed_form(oracle_lib):
    # hllc.cpp:527-537: phi = 0 for r <= 0, else min(2/(1+r), 2r/(1+r))
    for r, phi in [(-3.0, 0.0), (0.0, 0.0), (0.5, 2.0 / 3.0), (1.0, 1.0), (3.0, 0.5), (1e6, 2.0 / (1.0 + 1e6))]:
        assert lib().tjo_vanleer(r) == pytest.approx(phi, rel=1e-15)
    # this form is symmetric under r -> 1/r (the two branches swap)
    for r in (0.2, 0.7, 2.5, 9.0):
        assert lib().tjo_vanleer(r) == pytest.approx(lib().tjo_vanleer(1.0 / r), rel=1e-14)


def test_muscl_linear_data_exact_and_extremum_flat(oracle_lib):
    pb = one_cell_problem(n=12)
    L = pb.layout
    P = pb.prims
    x = np.arange(12, dtype=float)
    P[L.comp(C.ALPHA, 0)] = 0.5
    P[L.comp(C.ALPHA, 1)] = 0.5
    P[L.comp(C.RHO_K, 0)] = 1000.0 + 10.0 * x           # linear -> exact face values
    P[L.comp(C.RHO_K, 1)] = 50.0 + 0.0 * x
    P[L.comp(C.VELOCITY, 0, 0)] = 3.0 - (x - 6.0) ** 2   # extremum at x = 6 -> zero slope there
    P[L.comp(C.P)] = 1e5
    o = Oracle(pb)
    UL, UR = o.reconstruct(0)
    rk_l, rk_r = UL[L.comp(C.RHO_K, 0), 0, 0], UR[L.comp(C.RHO_K, 0), 0, 0]
    inner = slice(1, 11)
    assert np.allclose(rk_l[inner], 1000.0 + 10.0 * (x[inner] - 0.5), rtol=1e-15)
    assert np.allclose(rk_r[inner], 1000.0 + 10.0 * (x[inner] + 0.5), rtol=1e-15)
    u_l, u_r = UL[L.comp(C.VELOCITY, 0, 0), 0, 0], UR[L.comp(C.VELOCITY, 0, 0), 0, 0]
    assert u_l[6] == 3.0 and u_r[6] == 3.0
    # transmissive ends: foextrap ghost equals the edge cell -> zero slope
    assert rk_l[0] == rk_r[0] == 1000.0


# ---------------------------------------------------------------- EOS
def test_stiffened_gas_energy_and_sound(oracle_lib):
    g0, pi0, g1 = 4.4, 6e8, 1.4
    pb = one_cell_problem(n=3)
    L = pb.layout
    P = pb.prims
    a0 = np.array([1.0, 0.3, 1e-6])
    P[L.comp(C.ALPHA, 0), 0, 0] = a0
    P[L.comp(C.ALPHA, 1), 0, 0] = 1.0 - a0
    P[L.comp(C.RHO_K, 0)] = 1000.0
    P[L.comp(C.RHO_K, 1)] = 50.0
    P[L.comp(C.VELOCITY, 0, 0)] = 20.0
    P[L.comp(C.VELOCITY, 0, 1)] = -5.0
    p = 2e8
    P[L.comp(C.P)] = p
    o = Oracle(pb)
    o.sound()
    U = o.get()
    for i in range(3):
        a_0, a_1 = a0[i], 1.0 - a0[i]
        rho = a_0 * 1000.0 + a_1 * 50.0
        ke = 0.5 * rho * (20.0 ** 2 + 5.0 ** 2)
        # rho e = sum alpha_k (p + gamma_k pinf_k)/(gamma_k - 1)
        E = a_0 * (p + g0 * pi0) / (g0 - 1.0) + a_1 * p / (g1 - 1.0) + ke
        assert U[L.comp(C.TOTAL_E), 0, 0, i] == pytest.approx(E, rel=1e-14)
        # frozen mixture sound speed (boxaccesscellarray.cpp:254-285) with
        # c_k^2 = gamma_k (p + pinf_k)/rho_k
        c0sq, c1sq = g0 * (p + pi0) / 1000.0, g1 * p / 50.0
        num = (a_0 * 1000.0 / rho) * c0sq / (g0 - 1.0) + (a_1 * 50.0 / rho) * c1sq / (g1 - 1.0)
        den = a_0 / (g0 - 1.0) + a_1 / (g1 - 1.0)
        assert U[L.comp(C.SOUNDSPEED), 0, 0, i] == pytest.approx(math.sqrt(num / den), rel=1e-14)
    # single-material limit: c = sqrt(gamma (p + pinf)/rho)
    assert U[L.comp(C.SOUNDSPEED), 0, 0, 0] == pytest.approx(math.sqrt(g0 * (p + pi0) / 1000.0), rel=1e-14)


def test_c2p_inverts_p2c(oracle_lib):
    pb = C.shock_inclusion_2d(nx=32, ny=16, box=16, drift=(40.0, -30.0))
    o = Oracle(pb)
    before = o.get()
    o.c2p()
    after = o.get()
    L = pb.layout
    for c in L.prim_comps():
        scale = np.max(np.abs(before[c])) or 1.0
        assert np.max(np.abs(after[c] - before[c])) / scale < 1e-14


# ---------------------------------------------------------------- HLLC
def _cell(pb, a0, r0, r1, u, p):
    L = pb.layout
    v = np.zeros(L.ncomp)
    v[L.comp(C.ALPHA, 0)] = a0
    v[L.comp(C.ALPHA, 1)] = 1.0 - a0
    v[L.comp(C.RHO_K, 0)] = r0
    v[L.comp(C.RHO_K, 1)] = r1
    for q in range(3):
        v[L.comp(C.VELOCITY, 0, q)] = u[q]
    v[L.comp(C.P)] = p
    return v


def _physical_flux(pb, v, d, gam=(4.4, 1.4), pinf=(6e8, 0.0)):
    L = pb.layout
    a = [v[L.comp(C.ALPHA, m)] for m in range(2)]
    rk = [v[L.comp(C.RHO_K, m)] for m in range(2)]
    u = [v[L.comp(C.VELOCITY, 0, q)] for q in range(3)]
    p = v[L.comp(C.P)]
    rho = a[0] * rk[0] + a[1] * rk[1]
    E = sum(a[m] * (p + gam[m] * pinf[m]) / (gam[m] - 1.0) for m in range(2)) + 0.5 * rho * sum(x * x for x in u)
    f = np.zeros(L.ncomp)
    for m in range(2):
        f[L.comp(C.ALPHA, m)] = a[m] * u[d]
        f[L.comp(C.ALPHARHO, m)] = a[m] * rk[m] * u[d]
    for q in range(3):
        f[L.comp(C.RHOU, 0, q)] = rho * u[q] * u[d] + (p if q == d else 0.0)
    f[L.comp(C.TOTAL_E)] = (E + p) * u[d]
    return f


@pytest.mark.parametrize("d", [0, 1])
def test_hllc_consistency_identical_states(oracle_lib, d):
    pb = one_cell_problem(n=4, dim=2)
    pb.n = (4, 4, 1)
    pb.prims = np.zeros((pb.layout.ncomp,) + pb.shape)
    o = Oracle(pb)
    v = _cell(pb, 0.3, 1000.0, 50.0, (35.0, -12.0, 4.0), 3e8)
    F = o.face_flux(d, v, v)
    ref = _physical_flux(pb, v, d)
    for c in pb.layout.cons_comps():
        assert F[c] == pytest.approx(ref[c], rel=1e-12, abs=1e-9 * np.max(np.abs(ref)))


def test_hllc_stationary_contact(oracle_lib):
    pb = one_cell_problem(n=4)
    o = Oracle(pb)
    L = pb.layout
    left = _cell(pb, 1.0 - 1e-6, 1000.0, 50.0, (0.0, 0.0, 0.0), 1e6)
    right = _cell(pb, 1e-6, 1000.0, 50.0, (0.0, 0.0, 0.0), 1e6)
    F = o.face_flux(0, left, right)
    for m in range(2):
        assert F[L.comp(C.ALPHARHO, m)] == 0.0
        assert F[L.comp(C.ALPHA, m)] == 0.0
    assert F[L.comp(C.RHOU, 0, 0)] == pytest.approx(1e6, rel=1e-14)
    assert F[L.comp(C.TOTAL_E)] == 0.0
    assert F[L.comp(C.USTAR)] == 0.0


def test_hllc_mirror_symmetry(oracle_lib):
    pb = one_cell_problem(n=4)
    o = Oracle(pb)
    L = pb.layout
    a = _cell(pb, 0.7, 1000.0, 50.0, (120.0, 3.0, 0.0), 4e8)
    b = _cell(pb, 0.2, 1100.0, 40.0, (-60.0, 1.0, 0.0), 1e8)

    def mirror(v):
        w = v.copy()
        w[L.comp(C.VELOCITY, 0, 0)] *= -1.0
        return w

    F = o.face_flux(0, a, b)
    G = o.face_flux(0, mirror(b), mirror(a))
    for var in (C.ALPHARHO,):
        for m in range(2):
            c = L.comp(var, m)
            assert G[c] == pytest.approx(-F[c], rel=1e-12)
    assert G[L.comp(C.RHOU, 0, 0)] == pytest.approx(F[L.comp(C.RHOU, 0, 0)], rel=1e-12)
    assert G[L.comp(C.TOTAL_E)] == pytest.approx(-F[L.comp(C.TOTAL_E)], rel=1e-12)
    assert G[L.comp(C.USTAR)] == pytest.approx(-F[L.comp(C.USTAR)], rel=1e-12, abs=1e-12)


# ---------------------------------------------------------------- scheme invariants
def test_uniform_state_preserved(oracle_lib):
    pb = C.shock_inclusion_2d(nx=24, ny=16, box=8)
    L = pb.layout
    P = pb.prims
    P[L.comp(C.ALPHA, 0)] = 0.75
    P[L.comp(C.ALPHA, 1)] = 0.25
    P[L.comp(C.RHO_K, 0)] = 1.2
    P[L.comp(C.RHO_K, 1)] = 0.2
    P[L.comp(C.VELOCITY, 0, 0)] = 50.0
    P[L.comp(C.VELOCITY, 0, 1)] = -20.0
    P[L.comp(C.P)] = 1e5
    pb.bc_lo = pb.bc_hi = (0, 0, 0)
    o = Oracle(pb)
    s0 = o.get()
    for n in range(5):
        o.step(0.2 * o.timestep())
    s1 = o.get()
    for c in L.state_comps():
        assert np.max(np.abs(s1[c] - s0[c])) <= 1e-13 * max(np.max(np.abs(s0[c])), 1e-300) + 1e-300


def test_abgrall_pressure_velocity_equilibrium(oracle_lib):
    """Uniform p and u across a material interface stay uniform (the
    quasi-conservative alpha update keeps the mixture in equilibrium)."""
    pb = C.shock_tube_1d(n=200)
    L = pb.layout
    pb.prims[L.comp(C.P)] = 1e5
    pb.prims[L.comp(C.VELOCITY, 0, 0)] = 100.0
    o = Oracle(pb)
    for n in range(40):
        o.step(0.2 * o.timestep() if n < 5 else o.timestep())
    s = o.get()
    p = s[L.comp(C.P), 0, 0]
    u = s[L.comp(C.VELOCITY, 0, 0)]
    assert np.max(np.abs(p - 1e5)) / 1e5 < 1e-9
    assert np.max(np.abs(u - 100.0)) / 100.0 < 1e-9
    a0 = s[L.comp(C.ALPHA, 0), 0, 0]
    assert a0.min() < 0.5 < a0.max()  # the interface is still there (it moved, not vanished)


def inflow_balance(pb, s0, s1, t):
    """Expected change of the cell sums of alphaRho_m and E over time t for the
    shock/inclusion case while no wave has reached the boundaries: only the
    uniform post-shock inflow through the x = 0 face changes them (every other
    boundary face sees identical states at rest or a wall with zero normal
    velocity).  Returns [(measured, expected, total)] for alphaRho_0,
    alphaRho_1 and E, total being the cell sum at the start."""
    L = pb.layout
    rho2, u2, p2 = C.rankine_hugoniot(1.4, 1.0, 1e5, 1.22)
    a0 = 1.0 - C.ABSENT_ALPHA
    ny = pb.shape[1]
    dx = pb.L[0] / pb.n[0]
    rk1 = 0.138
    ar = [a0 * rho2, (1.0 - a0) * rk1]
    E_in = a0 * p2 / 0.4 + (1.0 - a0) * p2 / 0.67 + 0.5 * (ar[0] + ar[1]) * u2 * u2
    fl = [ar[0] * u2, ar[1] * u2, (E_in + p2) * u2]
    out = []
    for k, c in enumerate([L.comp(C.ALPHARHO, 0), L.comp(C.ALPHARHO, 1), L.comp(C.TOTAL_E)]):
        out.append((s1[c].sum() - s0[c].sum(), t / dx * ny * fl[k], s0[c].sum()))
    return out


def test_inflow_balance_of_mass_and_energy(oracle_lib):
    """Flux telescoping: interior faces cancel, so the cell sums change only by
    the boundary inflow (post-shock state entering at x = 0)."""
    pb = C.shock_inclusion_2d(nx=64, ny=32, box=16)
    o = Oracle(pb)
    s0 = o.get()
    t = 0.0
    for n in range(15):
        dt = 0.2 * o.timestep() if n < 5 else o.timestep()
        o.step(dt)
        t += dt
    s1 = o.get()
    for got, expect, total in inflow_balance(pb, s0, s1, t):
        assert abs(got - expect) <= 1e-13 * abs(total)


def test_two_material_riemann_problem_converges_to_exact(oracle_lib):
    from tests.exact_riemann import Side, sample, star

    left = Side(1000.0, 0.0, 1e9, 4.4, 6e8)
    right = Side(50.0, 0.0, 1e5, 1.4, 0.0)
    ps, us = star(left, right)
    errs = []
    for n in (100, 200, 400):
        pb = C.shock_tube_1d(n=n)
        o = Oracle(pb)
        steps, t = o.run(pb.final_time)
        assert t == pytest.approx(pb.final_time, rel=1e-12)
        s = o.get()
        L = pb.layout
        rho = s[L.comp(C.RHO), 0, 0]
        p = s[L.comp(C.P), 0, 0]
        x = (np.arange(n) + 0.5) / n
        re, ue, pe = sample(left, right, (x - 0.7) / t)
        errs.append(np.mean(np.abs(rho - re)) / np.mean(re))
        # the star-region plateau (between the contact and the shock)
        plateau = (x > 0.72) & (x < 0.78)
        assert np.max(np.abs(p[plateau] - ps)) / ps < 0.05
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 0.01


# ---------------------------------------------------------------- THINC
def test_thinc_face_values_preserve_cell_average(oracle_lib):
    """In 1D the THINC profile min + (range/2)(1 + theta tanh(beta(x - xt)))
    has its jump position xt fixed by the cell average; the reconstructed face
    values must be that profile at x = 0 and x = 1 (thinc.cpp:164-174)."""
    pb = one_cell_problem(n=5)
    pb.THINC = 1
    L = pb.layout
    P = pb.prims
    a = np.array([0.1, 0.1, 0.4, 0.9, 0.9])
    P[L.comp(C.ALPHA, 0), 0, 0] = a
    P[L.comp(C.ALPHA, 1), 0, 0] = 1.0 - a
    P[L.comp(C.RHO_K, 0)] = 1000.0
    P[L.comp(C.RHO_K, 1)] = 50.0
    P[L.comp(C.P)] = 1e5
    o = Oracle(pb)
    o.reconstruct(0)
    TL = o.get(Oracle.ULSTAR)[L.comp(C.ALPHA, 0), 0, 0, 2]
    TR = o.get(Oracle.URSTAR)[L.comp(C.ALPHA, 0), 0, 0, 2]
    mn, rng, beta = 0.1, 0.8, 1.6 + 0.01
    # solve the average condition for xt independently
    from scipy.optimize import brentq

    def avg(xt):
        return mn + 0.5 * rng * (1.0 + (math.log(math.cosh(beta * (1.0 - xt))) - math.log(math.cosh(beta * xt))) / beta)

    xt = brentq(lambda z: avg(z) - 0.4, -50.0, 50.0)
    assert TL == pytest.approx(mn + 0.5 * rng * (1.0 + math.tanh(-beta * xt)), rel=1e-9)
    assert TR == pytest.approx(mn + 0.5 * rng * (1.0 + math.tanh(beta * (1.0 - xt))), rel=1e-9)


# ---------------------------------------------------------------- clean-ups
def test_cleanup_alpha_cases(oracle_lib):
    pb = one_cell_problem(n=7)
    L = pb.layout
    cases = [(0.3, 0.6), (-0.1, 0.9), (1.2, 0.1), (0.4, np.nan), (np.nan, 0.7), (np.nan, np.nan), (0.5, 0.5)]
    P = pb.prims
    for i, (x, y) in enumerate(cases):
        P[L.comp(C.ALPHA, 0), 0, 0, i] = x
        P[L.comp(C.ALPHA, 1), 0, 0, i] = y
    o = Oracle(pb, init=False)
    o.set(P)
    assert o.cleanup_alpha() == 0
    s = o.get()
    got = [(s[L.comp(C.ALPHA, 0), 0, 0, i], s[L.comp(C.ALPHA, 1), 0, 0, i]) for i in range(len(cases))]
    expect = [(0.3 / 0.9, 0.6 / 0.9),
              (1e-6 / (1e-6 + 0.9), 0.9 / (1e-6 + 0.9)),
              (1.0 / 1.1, 0.1 / 1.1),
              (0.4, 0.6),          # NaN at index 1 -> 1 - sum
              (0.0, 1.0),          # NaN at index 0 stays 0 (boxaccesscellarray.cpp:553-556)
              (0.0, 1.0),          # two NaNs with the first at index 0 are not fatal
              (0.5, 0.5)]
    for g, e in zip(got, expect):
        assert g[0] == pytest.approx(e[0], rel=1e-14, abs=1e-300)
        assert g[1] == pytest.approx(e[1], rel=1e-14, abs=1e-300)


# ---------------------------------------------------------------- 3x3 algebra
def _dp(a):
    import ctypes
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def test_sqrtm_det_inverse(oracle_lib):
    rng = np.random.default_rng(7)
    for _ in range(50):
        M = rng.normal(size=(3, 3))
        A = np.ascontiguousarray(M @ M.T + 0.5 * np.eye(3))
        S = np.zeros((3, 3))
        lib().tjo_sqrtm(_dp(A), _dp(S))
        assert np.allclose(S, S.T, rtol=0, atol=1e-14 * np.abs(S).max())
        assert np.allclose(S @ S, A, rtol=1e-13, atol=1e-13 * np.abs(A).max())
        assert np.all(np.linalg.eigvalsh(S) > 0)
        Minv = np.zeros((3, 3))
        Mc = np.ascontiguousarray(M)
        lib().tjo_inv3(_dp(Mc), _dp(Minv))
        assert np.allclose(Minv @ M, np.eye(3), atol=1e-12)
        assert lib().tjo_det3(_dp(Mc)) == pytest.approx(np.linalg.det(M), rel=1e-12)
    D = np.ascontiguousarray(np.diag([4.0, 9.0, 0.25]))
    S = np.zeros((3, 3))
    lib().tjo_sqrtm(_dp(D), _dp(S))
    assert np.array_equal(S, np.diag([2.0, 3.0, 0.5]))


def test_maxsize_chops_like_amrex(oracle_lib):
    import ctypes

    def chop(n, chunk):
        out = (ctypes.c_int * 4096)()
        k = lib().tjo_chop(n, chunk, out)
        starts = list(out[:k])
        return [b - a for a, b in zip(starts, starts[1:] + [n])]

    assert chop(2048, 64) == [64] * 32
    assert chop(1000, 1 << 30) == [1000]
    s = chop(1000, 64)
    assert sum(s) == 1000 and max(s) <= 64 and len(s) == 16
    s = chop(300, 40)
    assert sum(s) == 300 and max(s) <= 40 and sorted(s) == s  # larger blocks at the high end


# ---------------------------------------------------------------- fixtures
def _fixture_cases():
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import make_fixtures

    return list(make_fixtures.cases())


@pytest.mark.parametrize("idx", [0, 1])
def test_oracle_reproduces_golden_fixtures(oracle_lib, idx):
    name, pb, nsteps = _fixture_cases()[idx]
    fx = np.load(os.path.join(ROOT, "tests", "golden", name + ".npz"))
    assert np.array_equal(fx["prims"], pb.prims)
    o = Oracle(pb)
    comps = list(fx["comps"])
    assert np.array_equal(o.get()[comps], fx["states"][0])
    for n in range(nsteps):
        dt = o.timestep()
        assert dt == fx["dts"][n]
        o.step(0.2 * dt if n < 5 else dt)
        assert np.array_equal(o.get()[comps], fx["states"][n + 1])


# ---------------------------------------------------------------- the C ABI library
def test_c_abi_library_exports_every_declared_symbol():
    import ctypes

    from tenjin_amd import LIB_PATH

    if not os.path.exists(LIB_PATH):
        pytest.skip("libtenjin.so not built")
    header = open(os.path.join(ROOT, "include", "tenjin.h")).read()
    declared = sorted(set(re.findall(r"\b(tj_[a-z_0-9]+)\s*\(", header)))
    assert len(declared) >= 20
    lib_ = ctypes.CDLL(LIB_PATH)
    missing = [s for s in declared if not hasattr(lib_, s)]
    assert not missing, missing


def test_peaks_library_loads_and_rejects_bad_arguments():
    """The device-peak kernels' library (bench.py's machine figures) loads
    without a device and refuses invalid arguments before any HIP call."""
    import ctypes

    path = os.path.join(ROOT, "tenjin_amd", "libtenjin_peaks.so")
    if not os.path.exists(path):
        pytest.skip("libtenjin_peaks.so not built")
    lib_ = ctypes.CDLL(path)
    assert hasattr(lib_, "tjp_hbm") and hasattr(lib_, "tjp_fma")
    lib_.tjp_hbm.argtypes = [ctypes.c_int, ctypes.c_longlong, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    lib_.tjp_fma.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    r, t = ctypes.c_double(), ctypes.c_double()
    assert lib_.tjp_hbm(3, 1 << 20, 1, ctypes.byref(r), ctypes.byref(t)) == 1  # no such kind
    assert lib_.tjp_hbm(0, 16, 1, ctypes.byref(r), ctypes.byref(t)) == 1  # smaller than one block's reads
    assert lib_.tjp_hbm(0, 1 << 20, 0, ctypes.byref(r), ctypes.byref(t)) == 1
    assert lib_.tjp_fma(1, 0, 1, ctypes.byref(r), ctypes.byref(t)) == 1
    assert lib_.tjp_fma(1, 8, 1, None, ctypes.byref(t)) == 1


def test_batch_create_refuses_more_problems_than_its_limit():
    """tj_batch_create refuses a batch of more than TJ_BATCH_MAX_PROBLEMS
    problems (the batch kernels take one launch row per problem) with
    TJ_EINVAL, before it touches a device."""
    import ctypes

    from tenjin_amd import LIB_PATH, LevelDesc, MFView

    if not os.path.exists(LIB_PATH):
        pytest.skip("libtenjin.so not built")
    header = open(os.path.join(ROOT, "include", "tenjin.h")).read()
    limit = int(re.search(r"#define TJ_BATCH_MAX_PROBLEMS (\d+)", header).group(1))
    einval = int(re.search(r"TJ_EINVAL = (-?\d+)", header).group(1))
    assert limit == 65535  # the grid's y extent
    lib_ = ctypes.CDLL(LIB_PATH)
    descs, S, T = (LevelDesc * 1)(), (MFView * 1)(), (MFView * 1)()
    out = ctypes.c_void_p(1)
    rc = lib_.tj_batch_create(ctypes.byref(out), 0, limit + 1, descs, S, T)
    assert rc == einval and not out.value


def test_bench_companion_child_errors_become_records():
    """A companion run that fails, prints no JSON line or overruns its time
    gives an error record in the bench line instead of ending the bench."""
    sys.path.insert(0, ROOT)
    import bench

    r = bench._child_line([sys.executable, "-c", "import sys; sys.stderr.write('boom'); sys.exit(3)"], 60)
    assert r["error"].startswith("exit status 3") and "boom" in r["error"]
    r = bench._child_line([sys.executable, "-c", "print('no json here')"], 60)
    assert r["error"].startswith("exit status 0")
    r = bench._child_line([sys.executable, "-c", "import time; time.sleep(5)"], 0.5)
    assert r["error"].startswith("timed out")
    r = bench._child_line([sys.executable, "-c", "print('x'); print('{\"a\": 1}')"], 60)
    assert r == {"a": 1}


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode != 0 and "one GPU" in r.stderr


def test_plotfile_roundtrip(tmp_path):
    """One-level AMReX plotfile writer and reader: uneven boxes in 2D and 3D,
    the header fields, the FAB encoding line and per-box min/max."""
    import os

    from tenjin_amd.plotfile import read_plotfile, write_plotfile

    rng = np.random.default_rng(5)
    for shape, boxes in (((1, 20, 33), [((0, 0, 0), (16, 9, 0)), ((17, 0, 0), (32, 9, 0)), ((0, 10, 0), (32, 19, 0))]),
                         ((6, 5, 4), [((0, 0, 0), (3, 4, 2)), ((0, 0, 3), (3, 4, 5))])):
        data = rng.standard_normal((3,) + shape)
        path = str(tmp_path / f"plt_{len(boxes)}")
        write_plotfile(path, data, ["a", "b", "c"], boxes, (1.0, 2.0, 3.0), 1.25e-3, 7)
        meta, back = read_plotfile(path)
        assert np.array_equal(back, data)
        assert meta["names"] == ["a", "b", "c"] and meta["step"] == 7 and meta["time"] == 1.25e-3
        assert meta["dim"] == (3 if shape[0] > 1 else 2)
        head = open(os.path.join(path, "Level_0", "Cell_D_00000"), "rb").readline().decode()
        assert head.startswith("FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))")
        cell_h = open(os.path.join(path, "Level_0", "Cell_H")).read().split("\n")
        assert cell_h[:4] == ["1", "1", "3", "0"]
