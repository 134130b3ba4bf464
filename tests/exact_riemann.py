"""Exact Riemann solver for two stiffened gases (Toro, Riemann Solvers and
Numerical Methods for Fluid Dynamics, ch. 4, with p replaced by p + p_inf in
each material's shock and rarefaction relations).  Test infrastructure."""
from __future__ import annotations

import math

import numpy as np


class Side:
    def __init__(self, rho, u, p, gamma, pinf):
        self.rho, self.u, self.p, self.g, self.pinf = rho, u, p, gamma, pinf
        self.c = math.sqrt(gamma * (p + pinf) / rho)


def _f(p, s: Side):
    """Pressure function and derivative (Toro 4.6-4.7 with p -> p + p_inf)."""
    g = s.g
    if p > s.p:  # shock
        A = 2.0 / ((g + 1.0) * s.rho)
        B = (g - 1.0) / (g + 1.0) * (s.p + s.pinf)
        q = math.sqrt(A / (p + s.pinf + B))
        return (p - s.p) * q, q * (1.0 - 0.5 * (p - s.p) / (p + s.pinf + B))
    # rarefaction
    pr = (p + s.pinf) / (s.p + s.pinf)
    f = 2.0 * s.c / (g - 1.0) * (pr ** ((g - 1.0) / (2.0 * g)) - 1.0)
    df = 1.0 / (s.rho * s.c) * pr ** (-(g + 1.0) / (2.0 * g))
    return f, df


def star(L: Side, R: Side):
    p = max(0.5 * (L.p + R.p), 1.0)
    for _ in range(200):
        fl, dl = _f(p, L)
        fr, dr = _f(p, R)
        res = fl + fr + (R.u - L.u)
        pn = p - res / (dl + dr)
        if pn <= -min(L.pinf, R.pinf):
            pn = 0.5 * (p - min(L.pinf, R.pinf))
        if abs(pn - p) <= 1e-14 * max(abs(p), 1.0):
            p = pn
            break
        p = pn
    fl, _ = _f(p, L)
    fr, _ = _f(p, R)
    u = 0.5 * (L.u + R.u) + 0.5 * (fr - fl)
    return p, u


def sample(L: Side, R: Side, xi: np.ndarray):
    """(rho, u, p) at similarity coordinates xi = (x - x0)/t."""
    ps, us = star(L, R)
    rho = np.empty_like(xi)
    u = np.empty_like(xi)
    p = np.empty_like(xi)
    for n, x in enumerate(xi):
        if x <= us:
            K, sgn = L, -1.0
        else:
            K, sgn = R, 1.0
        g, pi = K.g, K.pinf
        if ps > K.p:  # shock
            r = (ps + pi) / (K.p + pi)
            gr = (g - 1.0) / (g + 1.0)
            rho_s = K.rho * (r + gr) / (gr * r + 1.0)
            S = K.u + sgn * K.c * math.sqrt((g + 1.0) / (2.0 * g) * r + (g - 1.0) / (2.0 * g))
            behind = (x < S) if sgn < 0 else (x > S)
            if behind:
                rho[n], u[n], p[n] = K.rho, K.u, K.p
            else:
                rho[n], u[n], p[n] = rho_s, us, ps
        else:  # rarefaction
            rho_s = K.rho * ((ps + pi) / (K.p + pi)) ** (1.0 / g)
            c_s = K.c * ((ps + pi) / (K.p + pi)) ** ((g - 1.0) / (2.0 * g))
            head = K.u + sgn * K.c
            tail = us + sgn * c_s
            if (sgn < 0 and x <= head) or (sgn > 0 and x >= head):
                rho[n], u[n], p[n] = K.rho, K.u, K.p
            elif (sgn < 0 and x >= tail) or (sgn > 0 and x <= tail):
                rho[n], u[n], p[n] = rho_s, us, ps
            else:
                c = (2.0 / (g + 1.0) - sgn * (g - 1.0) / ((g + 1.0) * K.c) * (K.u - x)) * K.c
                uu = 2.0 / (g + 1.0) * (-sgn * K.c + (g - 1.0) / 2.0 * K.u + x)
                rho[n] = K.rho * (c / K.c) ** (2.0 / (g - 1.0))
                u[n] = uu
                p[n] = (K.p + pi) * (c / K.c) ** (2.0 * g / (g - 1.0)) - pi
    return rho, u, p
