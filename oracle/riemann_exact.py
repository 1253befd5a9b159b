"""ORACLE -- test infrastructure only.

Exact solution of the Riemann problem of the 1-D Euler equations for an ideal
gas with any ratio of specific heats, after E. F. Toro, *Riemann Solvers and
Numerical Methods for Fluid Dynamics*, 3rd ed., chapter 4 (pressure function
4.6-4.8, Newton iteration on it, and the wave-pattern sampling of 4.4).  The
reference's own Sod checker (``attic/GA.reproduce/riemann-solver.h:12``)
hard-codes gamma = 5/3, so this one is written from the textbook and takes
gamma as a parameter.

It pins the Hydro scheme (``examples/Hydro/HydroMain.hs``) on shock tubes:
the oracle's numerical solution must converge to ``sample``'s profile.
The solver itself is checked in the tests against Toro's published star
states and against the Rankine-Hugoniot and isentropic relations that the
waves it returns must satisfy.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class Side:
    rho: float
    u: float
    p: float


def _sound(s: Side, g: float) -> float:
    return math.sqrt(g * s.p / s.rho)


def _f(p: float, s: Side, g: float):
    """Toro's f_K(p) and its derivative: a shock branch above p_K, a rarefaction below."""
    c = _sound(s, g)
    if p > s.p:
        A = 2.0 / ((g + 1.0) * s.rho)
        B = (g - 1.0) / (g + 1.0) * s.p
        q = math.sqrt(A / (p + B))
        return (p - s.p) * q, q * (1.0 - 0.5 * (p - s.p) / (B + p))
    r = p / s.p
    return (2.0 * c / (g - 1.0) * (r ** ((g - 1.0) / (2.0 * g)) - 1.0),
            1.0 / (s.rho * c) * r ** (-(g + 1.0) / (2.0 * g)))


def star(L: Side, R: Side, g: float, tol: float = 1e-14):
    """Pressure and velocity of the star region (Newton on f_L + f_R + du = 0)."""
    cL, cR = _sound(L, g), _sound(R, g)
    if 2.0 * (cL + cR) / (g - 1.0) <= R.u - L.u:
        raise ValueError("vacuum is generated")
    # two-rarefaction estimate (Toro 4.46) as the starting point
    z = (g - 1.0) / (2.0 * g)
    p = ((cL + cR - 0.5 * (g - 1.0) * (R.u - L.u)) /
         (cL / L.p ** z + cR / R.p ** z)) ** (1.0 / z)
    p = max(p, 1e-12)
    for _ in range(100):
        fL, dL = _f(p, L, g)
        fR, dR = _f(p, R, g)
        p_new = max(p - (fL + fR + R.u - L.u) / (dL + dR), 1e-14)
        done = abs(p_new - p) <= tol * 0.5 * (p_new + p)
        p = p_new
        if done:
            break
    fL, _ = _f(p, L, g)
    fR, _ = _f(p, R, g)
    return p, 0.5 * (L.u + R.u) + 0.5 * (fR - fL)


def star_densities(L: Side, R: Side, g: float, p: float):
    def one(s):
        r = p / s.p
        if p > s.p:
            gm = (g - 1.0) / (g + 1.0)
            return s.rho * (r + gm) / (gm * r + 1.0)
        return s.rho * r ** (1.0 / g)
    return one(L), one(R)


def shock_speed(s: Side, g: float, p: float, right: bool) -> float:
    c = _sound(s, g)
    q = math.sqrt((g + 1.0) / (2.0 * g) * p / s.p + (g - 1.0) / (2.0 * g))
    return s.u + c * q if right else s.u - c * q


def sample(L: Side, R: Side, g: float, xi):
    """(rho, u, p) of the self-similar solution at xi = (x - x0) / t (array)."""
    xi = np.asarray(xi, dtype=np.float64)
    p, u = star(L, R, g)
    rL, rR = star_densities(L, R, g, p)
    out = np.empty((3,) + xi.shape)
    for idx, x in np.ndenumerate(xi):
        out[(slice(None),) + idx] = _sample_one(L, R, g, p, u, rL, rR, float(x))
    return out[0], out[1], out[2]


def _sample_one(L, R, g, ps, us, rsL, rsR, x):
    if x <= us:                                   # left of the contact
        s, sign, rs = L, 1.0, rsL
    else:
        s, sign, rs = R, -1.0, rsR
    c = _sound(s, g)
    if ps > s.p:                                  # shock
        S = shock_speed(s, g, ps, right=(sign < 0))
        behind = (x > S) if sign > 0 else (x < S)
        return (rs, us, ps) if behind else (s.rho, s.u, s.p)
    # rarefaction: head, tail
    cs = c * (ps / s.p) ** ((g - 1.0) / (2.0 * g))
    head = s.u - sign * c
    tail = us - sign * cs
    if (sign > 0 and x <= head) or (sign < 0 and x >= head):
        return s.rho, s.u, s.p
    if (sign > 0 and x >= tail) or (sign < 0 and x <= tail):
        return rs, us, ps
    k = 2.0 / (g + 1.0)
    base = k + sign * (g - 1.0) / ((g + 1.0) * c) * (s.u - x)
    rho = s.rho * base ** (2.0 / (g - 1.0))
    uu = k * (sign * c + 0.5 * (g - 1.0) * s.u + x)
    pp = s.p * base ** (2.0 * g / (g - 1.0))
    return rho, uu, pp
