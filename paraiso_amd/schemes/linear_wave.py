"""Linear wave equation in the reference's (f, g) form, in 1, 2 or 3 dimensions.

Restates ``examples-old/LinearWave/Convergence.hs:25-54``:

    f1 = f0 + dt * g0
    g1 = g0 + dt * c^2 / dx^2 * sum_k (f1[x - e_k] + f1[x + e_k] - 2 f1)
    energy = sum 0.5 * (c^2 * |grad f1|^2 + ((g0 + g1)/2)^2) * dx^d

(f = u, g = u_t; the pair is the leapfrog / symplectic-Euler update of
u_tt = c^2 Laplacian u with the 2d+1-point Laplacian).  The reference is 1-D
with c = 3.43, dx = 2 pi / n and dt = dx / c; here c, dx and dt are scalar
statics, so the benchmark sets c = 1, dx = 1 and its own Courant number.  The
energy reduction is accumulated in FP64.
"""
from __future__ import annotations

from ..om import builder as B
from ..om.arith import ReduceOp
from ..om.dyn import CType, Realm


def linear_wave_om(ndim: int = 3, real: CType = CType.Float, name: str = "LinearWave",
                   with_energy: bool = True):
    fieldF = B.static("fieldF", Realm.Array, real)
    fieldG = B.static("fieldG", Realm.Array, real)
    energy = B.static("energy", Realm.Scalar, CType.Double)
    c_s = B.static("c", Realm.Scalar, real)
    dx_s = B.static("dx", Realm.Scalar, real)
    dt_s = B.static("dt", Realm.Scalar, real)

    def unit(k, s):
        return [s if j == k else 0 for j in range(ndim)]

    def proceed():
        c = B.bind(B.broadcast(B.load(c_s)))
        dx = B.bind(B.broadcast(B.load(dx_s)))
        dt = B.bind(B.broadcast(B.load(dt_s)))
        f0 = B.bind(B.load(fieldF))
        g0 = B.bind(B.load(fieldG))
        f1 = B.bind(f0 + dt * g0)
        fR = [B.bind(B.shift(unit(k, -1), f1)) for k in range(ndim)]
        fL = [B.bind(B.shift(unit(k, 1), f1)) for k in range(ndim)]
        lap = B.bind(B.foldl1(lambda a, b: a + b,
                              [fL[k] + fR[k] - 2 * f1 for k in range(ndim)]))
        g1 = B.bind(g0 + dt * c ** 2 / dx ** 2 * lap)
        B.store(fieldF, f1)
        B.store(fieldG, g1)
        if with_energy:
            grad2 = B.foldl1(lambda a, b: a + b,
                             [((fR[k] - fL[k]) / (2 * dx)) ** 2 for k in range(ndim)])
            dens = 0.5 * (c ** 2 * grad2 + ((g0 + g1) / 2) ** 2) * dx ** ndim
            B.store(energy, B.reduce(ReduceOp.Sum, B.cast(dens, CType.Double)))

    def init():
        B.store(fieldF, 0)
        B.store(fieldG, 0)

    return B.make_om(name, [], [fieldF, fieldG, energy, c_s, dx_s, dt_s],
                     [("init", init), ("proceed", proceed)], ndim=ndim)
