"""ORACLE -- test infrastructure only.

An independent, vectorised restatement of the reference Hydro scheme
(``examples/Hydro/HydroMain.hs:134-276`` with the quantities of
``examples/Hydro/Hydro.hs:99-127``) in torch, written from the equations
rather than from the OM.  It runs on the CPU in float64 and serves as

* a second check of the OM-generated forward step (agreement to ~1e-13),
* the CPU FP64 reference adjoint: torch.autograd through the fixed-dt steps.

Arrays are (ny, nx); axis 0 of the OM (x) is torch dim 1, axis 1 (y) dim 0.
``shift v a`` of the OM (b[x] = a[x - v]) is ``torch.roll(a, v, dim)``.
Reference behaviours kept: the inverted HLLC pressure factor
(``hllcQ``, HydroMain.hs:264-265), star-region fluxes F(U*) (HydroMain.hs:266-276),
flux addition axis 1 then axis 0 (HydroMain.hs:175), floors max(1e-2, .) on
density and pressure (HydroMain.hs:177-180), std::max/std::min tie rules.
"""
from __future__ import annotations

import torch

DIM = {0: 1, 1: 0}


def shift(a, v_axis: int, axis: int):
    return torch.roll(a, shifts=v_axis, dims=DIM[axis])


def omax(a, b):        # std::max(a, b) = (a < b) ? b : a
    return torch.where(a < b, b, a)


def omin(a, b):        # std::min(a, b) = (b < a) ? b : a
    return torch.where(b < a, b, a)


class State:
    """A hydro state with every derived quantity computed from rho, v, p or rho, m, E."""

    def __init__(self, gamma, rho, v=None, p=None, m=None, e=None):
        g = gamma
        self.gamma = g
        self.rho = rho
        if v is not None:
            self.v = list(v)
            self.p = p
            self.m = [rho * vi for vi in self.v]
            self.ke = 0.5 * (self.v[0] * self.m[0] + self.v[1] * self.m[1])
            self.e = self.ke + self.p / (g - 1)
        else:
            self.m = list(m)
            self.e = e
            self.v = [mi / rho for mi in self.m]
            self.ke = 0.5 * (self.v[0] * self.m[0] + self.v[1] * self.m[1])
            self.p = (g - 1) * (self.e - self.ke)
        self.h = self.e + self.p
        self.c = torch.sqrt(g * self.p / rho)

    def dflux(self, i):
        return self.m[i]

    def mflux(self, j, i):
        return self.m[j] * self.v[i] + (self.p if i == j else 0.0 * self.p)

    def eflux(self, i):
        return self.h * self.v[i]


def interpolate(order, i, cell: State):
    def pair(x):
        a0, a1, a2, a3 = shift(x, 2, i), shift(x, 1, i), x, shift(x, -1, i)
        if order == 1:
            return a1, a2
        d01, d12, d23 = a1 - a0, a2 - a1, a3 - a2

        def absmaller(a, b):
            inner = torch.where(torch.abs(a) < torch.abs(b), a, b)
            return torch.where(a * b <= 0, torch.zeros_like(a), inner)
        return a1 + absmaller(d01, d12) / 2, a2 - absmaller(d12, d23) / 2
    rl, rr = pair(cell.rho)
    vl, vr = zip(*[pair(cell.v[k]) for k in range(2)])
    pl, pr = pair(cell.p)
    g = cell.gamma
    return State(g, rl, v=vl, p=pl), State(g, rr, v=vr, p=pr)


def hllc(i, L: State, R: State):
    g = L.gamma
    dens_mid = (L.rho + R.rho) / 2
    sound_mid = (L.c + R.c) / 2
    sl, sr = L.v[i], R.v[i]
    pstar = omax(torch.zeros_like(L.p), (L.p + R.p) / 2 - dens_mid * sound_mid * (sr - sl))

    def q(sp, p):
        return torch.where(p <= sp, torch.ones_like(p),
                           torch.sqrt(1 + (g + 1) / (2 * g) * (sp / p - 1)))
    shock_l = L.v[i] - L.c * q(pstar, L.p)
    shock_r = R.v[i] + R.c * q(pstar, R.p)
    shock_star = ((R.p - L.p + L.rho * sl * (shock_l - sl) - R.rho * sr * (shock_r - sr))
                  / (L.rho * (shock_l - sl) - R.rho * (shock_r - sr)))

    def star(sstar, shock, x):
        speed = x.v[i]
        dens = x.rho * (shock - speed) / (shock - sstar)
        mome = [dens * (sstar if j == i else x.v[j]) for j in range(2)]
        enrg = dens * (x.e / x.rho + (sstar - speed) * (sstar + x.p / x.rho / (shock - speed)))
        return State(g, dens, m=mome, e=enrg)
    ls, rs = star(shock_star, shock_l, L), star(shock_star, shock_r, R)
    c1, c2, c3 = 0 < shock_l, 0 < shock_star, 0 < shock_r

    def sel(f):
        return torch.where(c1, f(L), torch.where(c2, f(ls), torch.where(c3, f(rs), f(R))))
    return {"d": sel(lambda s: s.dflux(i)),
            "m": [sel(lambda s, j=j: s.mflux(j, i)) for j in range(2)],
            "e": sel(lambda s: s.eflux(i))}


def add_flux(dt, dR, wall, ex, cell: State) -> State:
    dtdx = dt / dR[ex]
    w = wall[ex]

    def right(a):
        return shift(a, -1, ex)
    dens = cell.rho + dtdx * (w["d"] - right(w["d"]))
    mome = [cell.m[j] + dtdx * (w["m"][j] - right(w["m"][j])) for j in range(2)]
    enrg = cell.e + dtdx * (w["e"] - right(w["e"]))
    return State(cell.gamma, dens, m=mome, e=enrg)


def proceed_single(order, dt, dR, cellF: State, cellS: State) -> State:
    wall = [hllc(i, *interpolate(order, i, cellF)) for i in range(2)]
    c = cellS
    for ex in (1, 0):
        c = add_flux(dt, dR, wall, ex, c)
    floor = torch.full_like(c.rho, 1e-2)
    return State(c.gamma, omax(floor, c.rho), v=c.v, p=omax(floor, c.p))


def cfl_dt(cell: State, dR, cfl):
    ts = [dR[i] / (cell.c + torch.abs(cell.v[i])) for i in range(2)]
    return cfl * torch.min(omin(ts[0], ts[1]))


def step(rho, v0, v1, p, dR, gamma, dt=None, cfl=None):
    """One ``proceed``; returns (rho, v0, v1, p, dt)."""
    cell = State(gamma, rho, v=(v0, v1), p=p)
    if dt is None:
        dt = cfl_dt(cell, dR, cfl)
    c2 = proceed_single(1, dt / 2, dR, cell, cell)
    c3 = proceed_single(2, dt, dR, c2, cell)
    return c3.rho, c3.v[0], c3.v[1], c3.p, dt


def run(state, nsteps, dR, gamma, dt=None, cfl=None):
    rho, v0, v1, p = state
    t = 0.0
    for _ in range(nsteps):
        rho, v0, v1, p, d = step(rho, v0, v1, p, dR, gamma, dt, cfl)
        t = t + d
    return (rho, v0, v1, p), t


def kinetic_energy_in_square(rho, v0, v1, lo=0.25, hi=0.75):
    """J = sum over cell centres in [lo, hi]^2 of rho |v|^2 / 2."""
    ny, nx = rho.shape
    x = (torch.arange(nx, dtype=rho.dtype) + 0.5) / nx
    y = (torch.arange(ny, dtype=rho.dtype) + 0.5) / ny
    mask = ((y >= lo) & (y <= hi))[:, None] & ((x >= lo) & (x <= hi))[None, :]
    return (0.5 * rho * (v0 * v0 + v1 * v1) * mask).sum(), mask
