"""Two-dimensional compressible Euler solver (HLLC, two-stage) as an OM program.

Restates ``examples/Hydro/Hydro.hs:37-181`` (the ``Hydrable`` quantities) and
``examples/Hydro/HydroMain.hs:44-318`` (``buildInit``, ``boundaryCondition``,
``buildProceed``, ``proceedSingle``, ``addFlux``, ``interpolate``, ``hllc``).

The reference builds every derived quantity as a fresh builder expression at
each use; the classes below keep that: ``Primitive``/``Conserved`` derive their
quantities from three stored fields each time they are asked, ``HydroRec``
holds bound values, and the boundary condition returns a record of *unbound*
Manifest-annotated selects that are re-run at each use -- exactly as the
Haskell ``(<$>)/(<*>)`` application does.  Node numbering and the per-node
expressions therefore match the reference's OM.

Parameters that the reference fixes in source and the benchmarks vary:
* ``real``: the content type (``Hydro.hs:37``),
* ``gamma``: the ratio of specific heats (``Hydro.hs:53-54`` hard-codes 5/3),
* ``dt_mode``: ``"cfl"`` (``HydroMain.hs:150-154``) or ``"fixed"`` (dt read
  from an extra scalar static ``dt``, for differentiable runs),
* ``legacy``: the older API of ``examples-old/Hydro-exampled/HydroMain.hs``
  (typed ``loadIndex``/``loadSize``, ``abs x = max x (negate x)``, no Manifest
  annotation on the HLLC selector),
* ``stages`` (fixed dt only): two more kernels that split ``proceed`` at its
  intermediate state (``cell2`` of ``buildProceed``, HydroMain.hs:156):
  ``proceed_stage1`` stores the half-step state into four more Array statics
  ``stage_*``, and ``proceed_stage2`` reads them and finishes the step.  Run in
  that order they store what ``proceed`` stores, with the same arithmetic; the
  adjoint is built per stage from them.
"""
from __future__ import annotations

from fractions import Fraction

from ..om import builder as B
from ..om.arith import ReduceOp
from ..om.dyn import CType, Realm

NDIM = 2


class _Ctx:
    def __init__(self, real: CType, gamma: Fraction, legacy: bool):
        self.real = real
        self.gamma = Fraction(gamma)
        self.legacy = legacy

    def k_gamma(self):
        return B.imm(self.gamma)

    def abs_(self, x):
        if self.legacy:
            return B.max_(x, -x)      # runs x twice, as the older front-end did
        return B.abs_(x)


def _delta(i, j):
    return B.imm(1 if i == j else 0)


class Hydrable:
    """Default methods of ``class Hydrable`` (Hydro.hs:99-127)."""

    def __init__(self, ctx: _Ctx):
        self.ctx = ctx

    def velocity(self, i):
        return self.momentum(i) / self.density()

    def pressure(self):
        return (self.ctx.k_gamma() - 1) * self.internal_energy()

    def momentum(self, i):
        return self.density() * self.velocity(i)

    def energy(self):
        return self.kinetic_energy() + 1 / (self.ctx.k_gamma() - 1) * self.pressure()

    def enthalpy(self):
        return self.energy() + self.pressure()

    def density_flux(self, i):
        return self.momentum(i)

    def momentum_flux(self, i, j):
        return self.momentum(i) * self.velocity(j) + self.pressure() * _delta(i, j)

    def energy_flux(self, i):
        return self.enthalpy() * self.velocity(i)

    def sound_speed(self):
        return sound_speed_(self.ctx, self.density(), self.pressure())

    def kinetic_energy(self):
        return 0.5 * B.contract(lambda i: self.velocity(i) * self.momentum(i), NDIM)

    def internal_energy(self):
        return self.energy() - self.kinetic_energy()


def sound_speed_(ctx, dens, pres):
    # Hydro.hs:57-58
    return B.sqrt(ctx.k_gamma() * pres / dens)


class Primitive(Hydrable):
    def __init__(self, ctx, dens, velo, pres):
        super().__init__(ctx)
        self._d, self._v, self._p = dens, list(velo), pres

    def density(self):
        return self._d

    def velocity(self, i):
        return self._v[i]

    def pressure(self):
        return self._p


class Conserved(Hydrable):
    def __init__(self, ctx, dens, mome, enrg):
        super().__init__(ctx)
        self._d, self._m, self._e = dens, list(mome), enrg

    def density(self):
        return self._d

    def momentum(self, i):
        return self._m[i]

    def energy(self):
        return self._e


FIELDS = ("density", "velocity", "pressure", "momentum", "energy", "enthalpy",
          "density_flux", "momentum_flux", "energy_flux", "sound_speed",
          "kinetic_energy", "internal_energy")
VEC_FIELDS = ("velocity", "momentum", "density_flux", "energy_flux")


def _field_slots():
    """Traversal order of the ``Hydro`` record (derived Traversable)."""
    for f in FIELDS:
        if f in VEC_FIELDS:
            for i in range(NDIM):
                yield (f, (i,))
        elif f == "momentum_flux":
            for i in range(NDIM):
                for j in range(NDIM):
                    yield (f, (i, j))
        else:
            yield (f, ())


SLOTS = list(_field_slots())


class HydroRec(Hydrable):
    """A ``Hydro a`` record: one builder per slot (bound or not)."""

    def __init__(self, ctx, slots: dict):
        super().__init__(ctx)
        self.slots = slots

    def density(self):
        return self.slots[("density", ())]

    def velocity(self, i):
        return self.slots[("velocity", (i,))]

    def pressure(self):
        return self.slots[("pressure", ())]

    def momentum(self, i):
        return self.slots[("momentum", (i,))]

    def energy(self):
        return self.slots[("energy", ())]

    def enthalpy(self):
        return self.slots[("enthalpy", ())]

    def density_flux(self, i):
        return self.slots[("density_flux", (i,))]

    def momentum_flux(self, i, j):
        return self.slots[("momentum_flux", (i, j))]

    def energy_flux(self, i):
        return self.slots[("energy_flux", (i,))]

    def sound_speed(self):
        return self.slots[("sound_speed", ())]

    def kinetic_energy(self):
        return self.slots[("kinetic_energy", ())]

    def internal_energy(self):
        return self.slots[("internal_energy", ())]

    def map_bind(self):
        """``mapM bind`` over the record."""
        return HydroRec(self.ctx, {s: B.bind(self.slots[s]) for s in SLOTS})


def _get(x: Hydrable, slot):
    f, idx = slot
    return getattr(x, f)(*idx)


def lift_a(ctx, f, *recs):
    """Applicative ``f <$> r1 <*> r2 ...`` over records: no nodes are built."""
    return HydroRec(ctx, {s: f(*[_get(r, s) for r in recs]) for s in SLOTS})


def bind_hydro(x: Hydrable) -> HydroRec:
    """``bindHydro`` (Hydro.hs:69-96): bind every derived quantity in order."""
    return HydroRec(x.ctx, {s: B.bind(_get(x, s)) for s in SLOTS})


def bind_primitive(ctx, d, v, p):
    return bind_hydro(Primitive(ctx, d, v, p))


def bind_conserved(ctx, d, m, e):
    return bind_hydro(Conserved(ctx, d, m, e))


# --------------------------------------------------------------------- kernels
class HydroProgram:
    def __init__(self, real=CType.Double, gamma=Fraction(5, 3), legacy=False,
                 dt_mode="cfl", name="Hydro", stages=False):
        self.ctx = _Ctx(real, gamma, legacy)
        self.real = real
        self.legacy = legacy
        self.dt_mode = dt_mode
        self.name = name
        R = real
        self.statics = (
            [B.static("generation", Realm.Scalar, CType.Int),
             B.static("time", Realm.Scalar, R),
             B.static("cfl", Realm.Scalar, R)]
            + [B.static("dR%d" % i, Realm.Scalar, R) for i in range(NDIM)]
            + [B.static("extent%d" % i, Realm.Scalar, R) for i in range(NDIM)]
            + [B.static("density", Realm.Array, R)]
            + [B.static("velocity%d" % i, Realm.Array, R) for i in range(NDIM)]
            + [B.static("pressure", Realm.Array, R)])
        if dt_mode == "fixed":
            self.statics.append(B.static("dt", Realm.Scalar, R))
        if stages and dt_mode != "fixed":
            raise ValueError("the two-stage kernels need a fixed dt")
        self.stages = stages
        if stages:
            self.statics += ([B.static("stage_density", Realm.Array, R)]
                             + [B.static("stage_velocity%d" % i, Realm.Array, R)
                                for i in range(NDIM)]
                             + [B.static("stage_pressure", Realm.Array, R)])
        self._by_name = {s.name: s for s in self.statics}

    def s(self, name):
        return self._by_name[name]

    def load_real(self, name):
        return B.load(self.s(name))

    def _icoord(self, axis):
        if self.legacy:
            return B.load_index(axis, self.real)
        return B.cast(B.load_index(axis), self.real)

    def _isize(self, axis):
        if self.legacy:
            return B.load_size(axis, self.real, Realm.Array)
        return B.broadcast(B.cast(B.load_size(axis), self.real))

    def _region(self, coord, extent):
        ex, ey = 0, 1
        # (cy > 0.47 ey) && ((cy < 0.53 ey) && (cx < 0)); && is infixr
        return coord[ey].gt(0.47 * extent[ey]) & (coord[ey].lt(0.53 * extent[ey]) & coord[ex].lt(0))

    def build_init(self):
        """``buildInit`` (HydroMain.hs:79-102)."""
        ctx = self.ctx
        dRG = [B.bind(self.load_real("dR%d" % i)) for i in range(NDIM)]
        extentG = [B.bind(self.load_real("extent%d" % i)) for i in range(NDIM)]
        dR = [B.bind(B.broadcast(x)) for x in dRG]
        extent = [B.bind(B.broadcast(x)) for x in extentG]
        icoord = [B.bind(self._icoord(ax)) for ax in range(NDIM)]
        coord = [B.bind(dR[i] * icoord[i]) for i in range(NDIM)]
        vplus = [B.imm(6), B.imm(0)]
        vminus = [B.imm(0), B.imm(0)]
        region = B.bind(self._region(coord, extent))
        velo = [B.bind(B.select(region, vplus[i], vminus[i])) for i in range(NDIM)]
        factor = B.bind(1 + 1e-3 * B.sin(6 * B.pi_() * coord[0]))
        B.store(self.s("density"), factor * ctx.k_gamma() * ctx.k_gamma() * B.select(region, 1, 10))
        for i in range(NDIM):
            B.store(self.s("velocity%d" % i), velo[i])
        B.store(self.s("pressure"), factor * 0.6)

    def boundary_condition(self, cell: HydroRec) -> HydroRec:
        """``boundaryCondition`` (HydroMain.hs:105-131)."""
        ctx = self.ctx
        dRG = [B.bind(self.load_real("dR%d" % i)) for i in range(NDIM)]
        extentG = [B.bind(self.load_real("extent%d" % i)) for i in range(NDIM)]
        dR = [B.bind(B.broadcast(x)) for x in dRG]
        extent = [B.bind(B.broadcast(x)) for x in extentG]
        icoord = [B.bind(self._icoord(ax)) for ax in range(NDIM)]
        isize = [B.bind(self._isize(ax)) for ax in range(NDIM)]
        coord = [B.bind(dR[i] * icoord[i]) for i in range(NDIM)]
        # the immediates carry the content type, which Haskell infers from the record's type
        # and which the selects below would otherwise learn only from the slots a kernel uses
        R = self.real
        vplus = [B.imm(6, R), B.imm(0, R)]
        vminus = [B.imm(0, R), B.imm(0, R)]
        region = B.bind(self._region(coord, extent))
        cell0 = bind_primitive(ctx,
                               ctx.k_gamma() * ctx.k_gamma()
                               * B.select(region, B.imm(1, R), B.imm(10, R)),
                               [B.select(region, vplus[i], vminus[i]) for i in range(NDIM)],
                               B.imm(0.6, R))
        out_of = B.bind(
            B.foldl1(lambda a, b: a | b, [icoord[i].lt(0) for i in range(NDIM)])
            | B.foldl1(lambda a, b: a | b, [icoord[i].ge(isize[i]) for i in range(NDIM)]))
        return lift_a(ctx, lambda a, b: B.manifest(B.select(out_of, a, b)), cell0, cell)

    def _step_inputs(self):
        """Loads, boundary condition and dt shared by ``proceed`` and its two stages."""
        ctx = self.ctx
        dens = B.bind(self.load_real("density"))
        velo = [B.bind(self.load_real("velocity%d" % i)) for i in range(NDIM)]
        pres = B.bind(self.load_real("pressure"))
        timeG = B.bind(self.load_real("time"))
        dRG = [B.bind(self.load_real("dR%d" % i)) for i in range(NDIM)]
        dR = [B.bind(B.broadcast(x)) for x in dRG]
        cell = self.boundary_condition(bind_primitive(ctx, dens, velo, pres))
        dtG = B.bind(self.load_real("dt"))
        dt = B.bind(B.broadcast(dtG))
        return cell, dR, dt, dtG, timeG

    def build_stage1(self):
        """The first half of ``buildProceed`` (HydroMain.hs:156): the order-1 half step
        from the state, stored into ``stage_*``."""
        cell, dR, dt, _, _ = self._step_inputs()
        cell2 = self.proceed_single(1, dt / 2, dR, cell, cell)
        B.store(self.s("stage_density"), cell2.density())
        for i in range(NDIM):
            B.store(self.s("stage_velocity%d" % i), cell2.velocity(i))
        B.store(self.s("stage_pressure"), cell2.pressure())

    def build_stage2(self):
        """The second half of ``buildProceed`` (HydroMain.hs:157-164): the order-2 step
        from the stored half-step state, applied to the state."""
        ctx = self.ctx
        cell, dR, dt, dtG, timeG = self._step_inputs()
        cell2 = bind_primitive(ctx, B.bind(self.load_real("stage_density")),
                               [B.bind(self.load_real("stage_velocity%d" % i))
                                for i in range(NDIM)],
                               B.bind(self.load_real("stage_pressure")))
        cell3 = self.proceed_single(2, dt, dR, cell2, cell)
        B.store(self.s("time"), timeG + dtG)
        B.store(self.s("density"), cell3.density())
        for i in range(NDIM):
            B.store(self.s("velocity%d" % i), cell3.velocity(i))
        B.store(self.s("pressure"), cell3.pressure())

    def build_proceed(self):
        """``buildProceed`` (HydroMain.hs:134-164)."""
        ctx = self.ctx
        dens = B.bind(self.load_real("density"))
        velo = [B.bind(self.load_real("velocity%d" % i)) for i in range(NDIM)]
        pres = B.bind(self.load_real("pressure"))
        timeG = B.bind(self.load_real("time"))
        cflG = B.bind(self.load_real("cfl"))
        dRG = [B.bind(self.load_real("dR%d" % i)) for i in range(NDIM)]
        dR = [B.bind(B.broadcast(x)) for x in dRG]
        cell0 = bind_primitive(ctx, dens, velo, pres)
        cell = self.boundary_condition(cell0)

        if self.dt_mode == "cfl":
            def timescale(i):
                return dR[i] / (cell.sound_speed() + ctx.abs_(cell.velocity(i)))
            dts = B.bind(B.foldl1(B.min_, [timescale(i) for i in range(NDIM)]))
            dtG = B.bind(cflG * B.reduce(ReduceOp.Min, dts))
        else:
            dtG = B.bind(self.load_real("dt"))
        dt = B.bind(B.broadcast(dtG))

        cell2 = self.proceed_single(1, dt / 2, dR, cell, cell)
        cell3 = self.proceed_single(2, dt, dR, cell2, cell)

        B.store(self.s("time"), timeG + dtG)
        B.store(self.s("density"), cell3.density())
        for i in range(NDIM):
            B.store(self.s("velocity%d" % i), cell3.velocity(i))
        B.store(self.s("pressure"), cell3.pressure())

    def proceed_single(self, order, dt, dR, cellF, cellS) -> HydroRec:
        """``proceedSingle`` (HydroMain.hs:169-180)."""
        ctx = self.ctx
        wall = []
        for i in range(NDIM):                    # sequence $ compose calcWall
            lp, rp = self.interpolate(order, i, cellF)
            wall.append(self.hllc(i, lp, rp))
        # foldl1 (.) over axes applied to cellS: the last axis is added first
        cellN = cellS
        for i in reversed(range(NDIM)):
            cellN = self.add_flux(dt, dR, wall, i, cellN)
        return bind_primitive(ctx, B.max_(1e-2, cellN.density()),
                              [cellN.velocity(i) for i in range(NDIM)],
                              B.max_(1e-2, cellN.pressure()))

    def add_flux(self, dt, dR, wall, ex, cell) -> HydroRec:
        """``addFlux`` (HydroMain.hs:185-196)."""
        ctx = self.ctx
        dtdx = B.bind(dt / dR[ex])
        left = wall[ex].map_bind()
        unit = [-(1 if k == ex else 0) for k in range(NDIM)]
        right = HydroRec(ctx, {s: B.bind(B.shift(unit, wall[ex].slots[s])) for s in SLOTS})
        dens1 = B.bind(cell.density() + dtdx * (left.density_flux(ex) - right.density_flux(ex)))
        mome1 = [B.bind(cell.momentum(j) + dtdx * (left.momentum_flux(j, ex) - right.momentum_flux(j, ex)))
                 for j in range(NDIM)]
        enrg1 = B.bind(cell.energy() + dtdx * (left.energy_flux(ex) - right.energy_flux(ex)))
        return bind_conserved(ctx, dens1, mome1, enrg1)

    def interpolate(self, order, i, cell):
        """``interpolate`` (HydroMain.hs:199-216)."""
        ctx = self.ctx

        def shifti(n):
            return [n if j == i else 0 for j in range(NDIM)]
        a = []
        for n in (2, 1, 0, -1):
            a.append(HydroRec(ctx, {s: B.bind(B.shift(shifti(n), cell.slots[s])) for s in SLOTS}))
        intp = {}
        for s in SLOTS:
            intp[s] = self.interpolate_single(order, *[a[k].slots[s] for k in range(4)])
        left = HydroRec(ctx, {s: intp[s][0] for s in SLOTS})
        right = HydroRec(ctx, {s: intp[s][1] for s in SLOTS})

        def bp(x):
            d1 = B.bind(x.density())
            v1 = [B.bind(x.velocity(k)) for k in range(NDIM)]
            p1 = B.bind(x.pressure())
            return bind_primitive(ctx, d1, v1, p1)
        return bp(left), bp(right)

    def interpolate_single(self, order, x0, x1, x2, x3):
        """``interpolateSingle`` (HydroMain.hs:218-235); runs when traversed."""
        if order == 1:
            return (x1, x2)
        ctx = self.ctx
        d01 = B.bind(x1 - x0)
        d12 = B.bind(x2 - x1)
        d23 = B.bind(x3 - x2)

        def absmaller(a, b):
            return B.select((a * b).le(0), 0, B.select(ctx.abs_(a).lt(ctx.abs_(b)), a, b))
        d1 = B.bind(absmaller(d01, d12))
        d2 = B.bind(absmaller(d12, d23))
        l = B.bind(x1 + d1 / 2)
        r = B.bind(x2 - d2 / 2)
        return (l, r)

    def hllc(self, i, left: HydroRec, right: HydroRec) -> HydroRec:
        """``hllc`` (HydroMain.hs:237-276)."""
        ctx = self.ctx
        kg = ctx.k_gamma
        dens_mid = B.bind((left.density() + right.density()) / 2)
        sound_mid = B.bind((left.sound_speed() + right.sound_speed()) / 2)
        speed_left = left.velocity(i)
        speed_right = right.velocity(i)
        pres_star = B.bind(B.max_(0, (left.pressure() + right.pressure()) / 2
                                  - dens_mid * sound_mid * (speed_right - speed_left)))

        def hllc_q(sp, p):
            return B.select(p.le(sp), 1, B.sqrt(1 + (kg() + 1) / (2 * kg()) * (sp / p - 1)))
        shock_left = B.bind(left.velocity(i) - left.sound_speed() * hllc_q(pres_star, left.pressure()))
        shock_right = B.bind(right.velocity(i) + right.sound_speed() * hllc_q(pres_star, right.pressure()))
        shock_star = B.bind(
            (right.pressure() - left.pressure()
             + left.density() * speed_left * (shock_left - speed_left)
             - right.density() * speed_right * (shock_right - speed_right))
            / (left.density() * (shock_left - speed_left)
               - right.density() * (shock_right - speed_right)))

        def star_state(star_shock, shock, x):
            speed = x.velocity(i)
            dens = B.bind(x.density() * (shock - speed) / (shock - star_shock))
            mome = [B.bind(dens * (star_shock if i == j else x.velocity(j))) for j in range(NDIM)]
            enrg = B.bind(dens * (x.energy() / x.density()
                                  + (star_shock - speed)
                                  * (star_shock + x.pressure() / x.density() / (shock - speed))))
            return bind_conserved(ctx, dens, mome, enrg)
        lesta = star_state(shock_star, shock_left, left)
        rista = star_state(shock_star, shock_right, right)

        def selector(a, b, c, d):
            sel = B.select(B.imm(0).lt(shock_left), a,
                           B.select(B.imm(0).lt(shock_star), b,
                                    B.select(B.imm(0).lt(shock_right), c, d)))
            return sel if ctx.legacy else B.manifest(sel)
        return lift_a(ctx, selector, left, lesta, rista, right).map_bind()

    def make(self):
        kernels = [("init", self.build_init), ("proceed", self.build_proceed)]
        if self.stages:
            kernels += [("proceed_stage1", self.build_stage1),
                        ("proceed_stage2", self.build_stage2)]
        return B.make_om(self.name, [], self.statics, kernels, ndim=NDIM)


def hydro_om(real=CType.Double, gamma=Fraction(5, 3), dt_mode="cfl", name="Hydro",
             stages=False):
    return HydroProgram(real, gamma, False, dt_mode, name, stages).make()


def hydro_om_legacy():
    """``examples-old/Hydro-exampled/HydroMain.hs`` (Float, gamma 5/3)."""
    return HydroProgram(CType.Float, Fraction(5, 3), True, "cfl", "Hydro").make()
