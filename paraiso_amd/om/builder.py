"""The Builder front-end: a Python counterpart of the reference's Builder monad.

Follows ``Language/Paraiso/OM/Builder/Internal.hs:53-484`` and
``OM/Builder/Boolean.hs:21-78``.

A ``Builder`` is a deferred action.  Running it appends nodes to the kernel
being built and returns a ``Value``.  As in the Haskell monad, an unbound
builder is re-run -- creating fresh nodes -- every time it is used, while
``bind(b)`` runs ``b`` once and returns a builder that only hands back the
resulting value.  Argument builders run left to right, then immediates are
materialised (``valueToNode``) left to right, then the instruction and its
result value are added.  This reproduces the reference's node numbering.

Kernels are built inside ``build_kernel`` (``Internal.hs:53-61``); the builder
functions use the kernel under construction implicitly.
"""
from __future__ import annotations

import math
import threading
from fractions import Fraction
from typing import Callable, List, Optional, Sequence

from .. import annotation as A
from .arith import COMPARISONS, Op, ReduceOp
from .dyn import CType, DynValue, Realm, TyVar, convert_literal
from .graph import (OM, Arith, Broadcast, Graph, Imm, Kernel, Load, LoadIndex,
                    LoadSize, Named, Node, Reduce, Setup, Shift, Store)

_tls = threading.local()


class BuilderState:
    def __init__(self, setup: Setup, ndim: int):
        self.setup = setup
        self.ndim = ndim
        self.graph = Graph()
        self.current_anot = list(setup.global_anot)


def _state() -> BuilderState:
    st = getattr(_tls, "state", None)
    if st is None:
        raise RuntimeError("builder used outside build_kernel")
    return st


class Value:
    """``FromNode`` (``node`` set) or ``FromImm`` (``literal`` set)."""
    __slots__ = ("node", "literal", "realm", "ctype")

    def __init__(self, realm: TyVar, ctype: TyVar, node: Optional[int] = None,
                 literal=None):
        self.node = node
        self.literal = literal
        self.realm = realm
        self.ctype = ctype


def _lit(x):
    """Normalise a Python literal; floats become exact decimal fractions.

    numpy scalars (np.float64 subclasses float) keep their exact binary value."""
    import numpy as _np
    if isinstance(x, _np.generic):
        if isinstance(x, _np.bool_):
            return bool(x)
        if isinstance(x, _np.integer):
            return int(x)
        return Fraction(float(x))
    if isinstance(x, bool):
        return x
    if isinstance(x, int):
        return x
    if isinstance(x, float):
        if x == math.pi:
            return _PI
        return Fraction(repr(x))
    if isinstance(x, (Fraction, _Pi)):
        return x
    raise TypeError("unsupported literal %r" % (x,))


class _Pi:
    """Haskell ``pi`` at the content type (correctly rounded per type)."""

    def __repr__(self):
        return "pi"


_PI = _Pi()


def _lit_ctype_hint(x) -> Optional[CType]:
    if isinstance(x, bool):
        return CType.Bool
    return None


class Builder:
    """A deferred OM action returning a ``Value``."""
    __slots__ = ("_run",)

    def __init__(self, run: Callable[[], Value]):
        self._run = run

    def run(self) -> Value:
        return self._run()

    # -- arithmetic (Internal.hs:323-418)
    def __add__(self, o):
        return mk_op2(Op.Add, self, lift(o))

    def __radd__(self, o):
        return mk_op2(Op.Add, lift(o), self)

    def __sub__(self, o):
        return mk_op2(Op.Sub, self, lift(o))

    def __rsub__(self, o):
        return mk_op2(Op.Sub, lift(o), self)

    def __mul__(self, o):
        return mk_op2(Op.Mul, self, lift(o))

    def __rmul__(self, o):
        return mk_op2(Op.Mul, lift(o), self)

    def __truediv__(self, o):
        return mk_op2(Op.Div, self, lift(o))

    def __rtruediv__(self, o):
        return mk_op2(Op.Div, lift(o), self)

    def __floordiv__(self, o):      # Haskell ``div`` -> Arith Div
        return mk_op2(Op.Div, self, lift(o))

    def __mod__(self, o):
        return mk_op2(Op.Mod, self, lift(o))

    def __neg__(self):
        return mk_op1(Op.Neg, self)

    def __pow__(self, n: int):
        return ipow(self, n)

    def pow(self, y):
        """``x ^/ y`` (Internal.hs:391): x to a rational power, ``A.Pow``."""
        return pow_(self, y)

    # -- Boolean (Internal.hs:381-386): & is &&, | is ||, ~ is not
    def __and__(self, o):
        return mk_op2(Op.And, self, lift(o))

    def __rand__(self, o):
        return mk_op2(Op.And, lift(o), self)

    def __or__(self, o):
        return mk_op2(Op.Or, self, lift(o))

    def __ror__(self, o):
        return mk_op2(Op.Or, lift(o), self)

    def __invert__(self):
        return mk_op1(Op.Not, self)

    # -- comparisons (Boolean.hs:41-58)
    def eq(self, o):
        return mk_op2b(Op.EQ, self, lift(o))

    def ne(self, o):
        return mk_op2b(Op.NE, self, lift(o))

    def lt(self, o):
        return mk_op2b(Op.LT, self, lift(o))

    def le(self, o):
        return mk_op2b(Op.LE, self, lift(o))

    def gt(self, o):
        return mk_op2b(Op.GT, self, lift(o))

    def ge(self, o):
        return mk_op2b(Op.GE, self, lift(o))

    def __bool__(self):
        raise TypeError("a Builder has no truth value; use select/eq/...")


def lift(x) -> Builder:
    if isinstance(x, Builder):
        return x
    return imm(x)


def _ret(v: Value) -> Builder:
    return Builder(lambda: v)


# ---------------------------------------------------------------- graph plumbing
def add_node(froms: Sequence[int], node: Node) -> int:
    return _state().graph.add(list(froms), node)


def add_node_e(froms: Sequence[int], *, dyn_ty=None, inst=None) -> int:
    st = _state()
    node = Node(inst=inst, anot=list(st.current_anot), ty=dyn_ty)
    return add_node(froms, node)


def value_to_node(v: Value) -> int:
    """``valueToNode`` (Internal.hs:132-142): immediates become Imm + value nodes."""
    if v.node is not None:
        return v.node
    n0 = add_node_e([], inst=_PendingImm(v.literal, v.ctype), dyn_ty=(v.realm, v.ctype))
    n1 = add_node_e([n0], dyn_ty=(v.realm, v.ctype))
    return n1


class _PendingImm:
    """Placeholder instruction for an immediate whose type is not yet known."""
    __slots__ = ("literal", "ctype")

    def __init__(self, literal, ctype: TyVar):
        self.literal = literal
        self.ctype = ctype


def _new_value(froms, realm: TyVar, ctype: TyVar, inst) -> Value:
    n0 = add_node_e(froms, inst=inst)
    n1 = add_node_e([n0], dyn_ty=(realm, ctype))
    return Value(realm, ctype, node=n1)


# ---------------------------------------------------------------- primitives
def bind(b: Builder) -> Builder:
    """Run ``b`` now; the returned builder yields its value without new nodes."""
    v = lift(b).run()
    return _ret(v)


def imm(x, ctype: Optional[CType] = None, realm: Optional[Realm] = None) -> Builder:
    lit = _lit(x)

    def run():
        return Value(TyVar(realm), TyVar(ctype or _lit_ctype_hint(lit)), literal=lit)
    return Builder(run)


def _named(static) -> Named:
    if isinstance(static, Named):
        return static
    raise TypeError("expected a Named static, got %r" % (static,))


def _lookup_static(name: str, dyn_realm: TyVar, dyn_ctype: TyVar) -> int:
    """``lookUpStatic`` (Internal.hs:146-159)."""
    st = _state()
    hits = [i for i, s in enumerate(st.setup.static_values) if s.name == name]
    if len(hits) != 1:
        raise KeyError("%d match found for '%s'" % (len(hits), name))
    idx = hits[0]
    dv = st.setup.static_values[idx].dyn
    dyn_realm.unify(TyVar(dv.realm), "static '%s' realm" % name)
    dyn_ctype.unify(TyVar(dv.ctype), "static '%s' type" % name)
    return idx


def load(static) -> Builder:
    nv = _named(static)

    def run():
        r, c = TyVar(nv.dyn.realm), TyVar(nv.dyn.ctype)
        idx = _lookup_static(nv.name, r, c)
        return _new_value([], r, c, Load(idx))
    return Builder(run)


def store(static, b) -> Builder:
    """Store a value to a static (Internal.hs:182-194); runs immediately."""
    nv = _named(static)
    v = lift(b).run()
    idx = _lookup_static(nv.name, v.realm, v.ctype)
    n0 = value_to_node(v)
    add_node_e([n0], inst=Store(idx))
    return None


def reduce(op: ReduceOp, b) -> Builder:
    def run():
        v1 = lift(b).run()
        v1.realm.unify(TyVar(Realm.Array), "reduce argument realm")
        n1 = value_to_node(v1)
        return _new_value([n1], TyVar(Realm.Scalar), v1.ctype, Reduce(op))
    return Builder(run)


def broadcast(b) -> Builder:
    def run():
        v1 = lift(b).run()
        v1.realm.unify(TyVar(Realm.Scalar), "broadcast argument realm")
        n1 = value_to_node(v1)
        return _new_value([n1], TyVar(Realm.Array), v1.ctype, Broadcast())
    return Builder(run)


def load_index(axis: int, ctype: CType = CType.Int) -> Builder:
    """``loadIndex`` (Internal.hs:231-241); ``ctype`` is the gauge type."""
    def run():
        return _new_value([], TyVar(Realm.Array), TyVar(ctype), LoadIndex(axis))
    return Builder(run)


def load_size(axis: int, ctype: CType = CType.Int, realm: Realm = Realm.Scalar) -> Builder:
    """``loadSize`` (Internal.hs:244-254).  The older API let the caller pick
    the realm and content type, which ``realm``/``ctype`` reproduce."""
    def run():
        return _new_value([], TyVar(realm), TyVar(ctype), LoadSize(axis))
    return Builder(run)


def shift(vec: Sequence[int], b) -> Builder:
    vec = tuple(int(x) for x in vec)

    def run():
        v1 = lift(b).run()
        v1.realm.unify(TyVar(Realm.Array), "shift argument realm")
        n1 = value_to_node(v1)
        return _new_value([n1], v1.realm, v1.ctype, Shift(vec))
    return Builder(run)


def mk_op1(op: Op, b) -> Builder:
    def run():
        v1 = lift(b).run()
        n1 = value_to_node(v1)
        return _new_value([n1], v1.realm, v1.ctype, Arith(op))
    return Builder(run)


def mk_op2(op: Op, b1, b2) -> Builder:
    def run():
        v1 = lift(b1).run()
        v2 = lift(b2).run()
        v1.realm.unify(v2.realm, "realm of %s" % op.value)
        v1.ctype.unify(v2.ctype, "type of %s" % op.value)
        n1 = value_to_node(v1)
        n2 = value_to_node(v2)
        return _new_value([n1, n2], v1.realm, v1.ctype, Arith(op))
    return Builder(run)


def mk_op2b(op: Op, b1, b2) -> Builder:
    """Comparison returning Bool (Boolean.hs:21-35)."""
    assert op in COMPARISONS

    def run():
        v1 = lift(b1).run()
        v2 = lift(b2).run()
        v1.realm.unify(v2.realm, "realm of %s" % op.value)
        v1.ctype.unify(v2.ctype, "type of %s" % op.value)
        n1 = value_to_node(v1)
        n2 = value_to_node(v2)
        return _new_value([n1, n2], v1.realm, TyVar(CType.Bool), Arith(op))
    return Builder(run)


def select(bb, b1, b2) -> Builder:
    """``select`` (Boolean.hs:61-78)."""
    def run():
        vb = lift(bb).run()
        v1 = lift(b1).run()
        v2 = lift(b2).run()
        vb.ctype.unify(TyVar(CType.Bool), "select condition")
        vb.realm.unify(v1.realm, "select realm")
        v1.realm.unify(v2.realm, "select realm")
        v1.ctype.unify(v2.ctype, "select branches")
        nb = value_to_node(vb)
        n1 = value_to_node(v1)
        n2 = value_to_node(v2)
        return _new_value([nb, n1, n2], v1.realm, v1.ctype, Arith(Op.Select))
    return Builder(run)


def cast(b, to: CType) -> Builder:
    """``cast``/``castTo`` (Internal.hs:422-446)."""
    def run():
        v1 = lift(b).run()
        n1 = value_to_node(v1)
        return _new_value([n1], v1.realm, TyVar(to), Arith(Op.Cast, cast_to=to))
    return Builder(run)


def annotate(f: Callable[[list], list], b) -> Builder:
    """``annotate`` / ``<?>`` (Internal.hs:465-484)."""
    def run():
        v1 = lift(b).run()
        n1 = value_to_node(v1)
        node = _state().graph.nodes[n1]
        node.anot = f(node.anot)
        return Value(v1.realm, v1.ctype, node=n1)
    return Builder(run)


def with_annotation(f: Callable[[list], list], b) -> Builder:
    def run():
        st = _state()
        cur0 = st.current_anot
        st.current_anot = f(list(cur0))
        try:
            ret = lift(b).run()
        finally:
            st.current_anot = cur0
        return ret
    return Builder(run)


def manifest(b) -> Builder:
    """``Anot.add Alloc.Manifest <?> b``."""
    return annotate(lambda a: A.add(A.Allocation.Manifest, a), b)


# ---- numeric class methods of the Builder (Internal.hs:323-418)
def sqrt(b):
    return mk_op1(Op.Sqrt, b)


def abs_(b):
    return mk_op1(Op.Abs, b)


def signum(b):
    return mk_op1(Op.Signum, b)


def recip(b):
    return mk_op1(Op.Inv, b)


def exp(b):
    return mk_op1(Op.Exp, b)


def log(b):
    return mk_op1(Op.Log, b)


def sin(b):
    return mk_op1(Op.Sin, b)


def cos(b):
    return mk_op1(Op.Cos, b)


def tan(b):
    return mk_op1(Op.Tan, b)


def asin(b):
    return mk_op1(Op.Asin, b)


def acos(b):
    return mk_op1(Op.Acos, b)


def atan(b):
    return mk_op1(Op.Atan, b)


def pow_(x, y) -> Builder:
    """``x ^/ y = mkOp2 A.Pow x (fromRational' y)`` (Internal.hs:391): the exponent is a
    rational literal, made an immediate of x's type."""
    if isinstance(y, Builder):
        raise TypeError("^/ takes a rational exponent, not a builder")
    return mk_op2(Op.Pow, lift(x), imm(y))


def atan2(y, x) -> Builder:
    """``A.Atan2`` (Arithmetic.hs:60), which the CPlusPlus emitter prints as ``atan2``
    (PlanTrans.hs:702)."""
    return mk_op2(Op.Atan2, lift(y), lift(x))


def max_(a, b):
    return mk_op2(Op.Max, lift(a), lift(b))


def min_(a, b):
    return mk_op2(Op.Min, lift(a), lift(b))


def pi_() -> Builder:
    return imm(_PI)


def zero() -> Builder:
    return imm(0)


def ipow(a, n: int) -> Builder:
    """Ring ``a ^ n`` by repeated squaring (Internal.hs:335-347)."""
    a = lift(a)
    if n == 0:
        return imm(1)
    if n == 1:
        return a

    def run():
        ba = bind(a)

        def f(x, n2):
            if n2 == 1:
                return x
            n3 = n2 // 2
            bx_n3 = bind(f(x, n3))
            sq = bx_n3 * bx_n3
            return (x * sq) if n2 - 2 * n3 > 0 else sq
        return f(ba, n).run()
    return Builder(run)


def foldl1(f, xs: Sequence):
    xs = list(xs)
    acc = xs[0]
    for x in xs[1:]:
        acc = f(acc, x)
    return acc


def foldl(f, z, xs: Sequence):
    acc = z
    for x in xs:
        acc = f(acc, x)
    return acc


def sum_(xs: Sequence):
    """NumericPrelude ``sum = foldl (+) zero``."""
    return foldl(lambda a, b: lift(a) + b, zero(), xs)


def contract(f, ndim: int):
    """typelevel-tensor ``contract f = foldl (+) zero (compose f)``."""
    return foldl(lambda a, b: lift(a) + b, zero(), [f(i) for i in range(ndim)])


# ---------------------------------------------------------------- kernels and machines
def _resolve(graph: Graph) -> None:
    for i, node in enumerate(graph.nodes):
        if node.ty is not None:
            r, c = node.ty[0].get(), node.ty[1].get()
            if r is None or c is None:
                raise TypeError("ambiguous OM type at node %d (realm=%s, type=%s)"
                                % (i, r, c))
            if node.inst is None:
                node.dyn = DynValue(r, c)
            elif isinstance(node.inst, _PendingImm):
                lit = node.inst.literal
                if isinstance(lit, _Pi):
                    lit = Fraction(math.pi) if c is CType.Double else Fraction(math.pi)
                exact = lit if c.is_float and isinstance(lit, (int, Fraction)) else None
                node.inst = Imm(convert_literal(lit, c), c,
                                None if exact is None else Fraction(exact))
            node.ty = None
        elif isinstance(node.inst, _PendingImm):
            raise TypeError("untyped immediate at node %d" % i)


def build_kernel(setup: Setup, name: str, body: Callable[[], None], ndim: int) -> Kernel:
    """``buildKernel`` (Internal.hs:53-61): run ``body`` against a fresh graph."""
    prev = getattr(_tls, "state", None)
    st = BuilderState(setup, ndim)
    _tls.state = st
    try:
        body()
    finally:
        _tls.state = prev
    _resolve(st.graph)
    return Kernel(name, st.graph)


def make_om(name: str, global_anot: list, statics: List[Named],
            kernels: List[tuple], ndim: int) -> OM:
    """``makeOM`` (OM.hs:28-41)."""
    setup = Setup(list(statics), list(global_anot))
    ks = [build_kernel(setup, kname, body, ndim) for kname, body in kernels]
    return OM(name, setup, ks, ndim)


def static(name: str, realm: Realm, ctype: CType) -> Named:
    return Named(name, DynValue(realm, ctype))
