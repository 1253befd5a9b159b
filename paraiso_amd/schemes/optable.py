"""Every operator the OM can lower, one per Array static: a program that checks the
operator table of the emitters (``OM/Arithmetic.hs:20-110``, printed by
``Generator/PlanTrans.hs:670-711``) on the GPU against the oracle.

Inputs are three Array statics with values in every operator's domain:
``x`` in (-0.95, 0.95) (the inverse trigonometric functions), ``y`` in
(0.01, 8) (log, sqrt, a rational power) and ``w`` in (-1, 1) (atan2's second
argument, signs of both kinds).  Integer tables use ``i`` and a nonzero ``j``.
``Signum`` is in its own kernel: the reference prints it as a call to an
undefined ``signum`` (PlanTrans.hs:706), so no oracle compiles it.
"""
from __future__ import annotations

from fractions import Fraction

from ..om import builder as B
from ..om.dyn import CType, Realm

POW_EXPONENT = Fraction(17, 10)
FLOAT_OPS = ["add", "sub", "mul", "div", "neg", "inv", "max", "min", "abs", "select",
             "eq", "ne", "lt", "le", "gt", "ge", "and", "or", "not",
             "sqrt", "exp", "log", "sin", "cos", "tan", "asin", "acos", "atan", "atan2",
             "pow", "cast_int", "cast_other"]
INT_OPS = ["add", "sub", "mul", "div", "mod", "neg", "max", "min", "abs", "select",
           "eq", "ne", "lt", "le", "gt", "ge", "cast_other"]
BOOL_RESULTS = {"eq", "ne", "lt", "le", "gt", "ge", "and", "or", "not"}
TRANSCENDENTAL_OPS = {"exp", "log", "sin", "cos", "tan", "asin", "acos", "atan", "atan2", "pow"}


def out_type(op: str, real: CType) -> CType:
    if op in BOOL_RESULTS:
        return CType.Int       # stored through select(b, 1, 0): vector<bool> has no data()
    if op == "cast_int":
        return CType.Int
    if op == "cast_other":
        return {CType.Float: CType.Double, CType.Double: CType.Float,
                CType.Int: CType.Double}[real]
    return real


def optable_om(real: CType, signum: bool = True):
    ops = INT_OPS if real is CType.Int else FLOAT_OPS
    ins = ["i", "j"] if real is CType.Int else ["x", "y", "w"]
    statics = [B.static(n, Realm.Array, real) for n in ins]
    statics += [B.static("out_" + op, Realm.Array, out_type(op, real)) for op in ops]
    if real is not CType.Int and signum:
        statics.append(B.static("out_signum", Realm.Array, real))
    by = {s.name: s for s in statics}

    def apply():
        if real is CType.Int:
            a, b = B.bind(B.load(by["i"])), B.bind(B.load(by["j"]))
        else:
            a, b, w = (B.bind(B.load(by[n])) for n in ins)
        f = {
            "add": lambda: a + b, "sub": lambda: a - b, "mul": lambda: a * b,
            "div": lambda: a / b, "mod": lambda: a % b, "neg": lambda: -a,
            "inv": lambda: B.recip(b), "max": lambda: B.max_(a, b), "min": lambda: B.min_(a, b),
            "abs": lambda: B.abs_(a), "select": lambda: B.select(a.lt(b), a, b),
            "eq": lambda: a.eq(b), "ne": lambda: a.ne(b), "lt": lambda: a.lt(b),
            "le": lambda: a.le(b), "gt": lambda: a.gt(b), "ge": lambda: a.ge(b),
            "and": lambda: a.lt(0) & b.gt(4), "or": lambda: a.lt(0) | b.gt(4),
            "not": lambda: ~a.lt(0),
            "sqrt": lambda: B.sqrt(b), "exp": lambda: B.exp(b), "log": lambda: B.log(b),
            "sin": lambda: B.sin(b), "cos": lambda: B.cos(b), "tan": lambda: B.tan(a),
            "asin": lambda: B.asin(a), "acos": lambda: B.acos(a), "atan": lambda: B.atan(b),
            "atan2": lambda: B.atan2(a, w), "pow": lambda: B.pow_(b, POW_EXPONENT),
            "cast_int": lambda: B.cast(b * 100, CType.Int),
            "cast_other": lambda: B.cast(a, out_type("cast_other", real)),
        }
        one, zero = B.imm(1, CType.Int), B.imm(0, CType.Int)
        for op in ops:
            v = f[op]()
            if op in BOOL_RESULTS:
                v = B.select(v, one, zero)
            B.store(by["out_" + op], v)

    def signum_kernel():
        B.store(by["out_signum"], B.signum(B.load(by["x"])))

    kernels = [("apply", apply)]
    if real is not CType.Int and signum:
        kernels.append(("signum", signum_kernel))
    return B.make_om("OpTable", [], statics, kernels, ndim=2)


def inputs(real: CType, shape, seed: int):
    """Seeded inputs in the operators' domains; shape (ny, nx)."""
    import numpy as np
    from ..inputs import uniform01
    n = int(np.prod(shape))
    u = [uniform01(seed, n, stream=k).reshape(shape) for k in range(3)]
    if real is CType.Int:
        i = np.rint(u[0] * 2000 - 1000).astype(np.int32)
        j = np.rint(u[1] * 60 - 30).astype(np.int32)
        j[j == 0] = 7
        return {"i": i, "j": j}
    dt = real.np_dtype
    x = (u[0] * 1.9 - 0.95).astype(dt)
    x.flat[::17] = 0                      # exact zeros for abs, signum and the comparisons
    y = (0.01 + 7.99 * u[1]).astype(dt)
    y.flat[5::23] = np.abs(x.flat[5::23]) + dt(0.01)    # some equal pairs x == y in (0, 1)
    x.flat[5::23] = y.flat[5::23]
    w = (u[2] * 2 - 1).astype(dt)
    return {"x": x, "y": y, "w": w}
