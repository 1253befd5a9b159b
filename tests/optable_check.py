"""Shared checks of the operator-table program (paraiso_amd/schemes/optable.py): the
closed form of every operator in numpy, and distances in units in the last place."""
import numpy as np

from paraiso_amd.om.dyn import CType
from paraiso_amd.schemes import optable as T

TRANSCENDENTAL = T.TRANSCENDENTAL_OPS


def ulp_distance(a, b):
    """|a - b| counted in representable values of their type (exact integers)."""
    a, b = np.asarray(a), np.asarray(b)
    it = {np.dtype(np.float32): np.int32, np.dtype(np.float64): np.int64}[a.dtype]
    lo = int(np.iinfo(it).min)

    def ordered(x):
        i = x.view(it).astype(object)
        return np.where(i < 0, lo - i, i)
    return np.abs(ordered(a) - ordered(b)).astype(np.float64)


def closed_form(op, real, ins):
    """The operator's value by its definition: IEEE arithmetic in the operand type, C's
    truncating integer division and remainder, std::max/std::min, libm in double."""
    if real is CType.Int:
        a, b = ins["i"].astype(np.int64), ins["j"].astype(np.int64)
    else:
        a, b, w = ins["x"], ins["y"], ins["w"]
    dt = T.out_type(op, real).np_dtype
    with np.errstate(all="ignore"):
        if op in T.TRANSCENDENTAL_OPS:
            f = {"exp": lambda: np.exp(b.astype(np.float64)),
                 "log": lambda: np.log(b.astype(np.float64)),
                 "sin": lambda: np.sin(b.astype(np.float64)),
                 "cos": lambda: np.cos(b.astype(np.float64)),
                 "tan": lambda: np.tan(a.astype(np.float64)),
                 "asin": lambda: np.arcsin(a.astype(np.float64)),
                 "acos": lambda: np.arccos(a.astype(np.float64)),
                 "atan": lambda: np.arctan(b.astype(np.float64)),
                 "atan2": lambda: np.arctan2(a.astype(np.float64), w.astype(np.float64)),
                 "pow": lambda: np.power(b.astype(np.float64),
                                         float(np.asarray(float(T.POW_EXPONENT), dtype=real.np_dtype)))}
            return f[op]().astype(dt)
        r = {"add": lambda: a + b, "sub": lambda: a - b, "mul": lambda: a * b,
             "div": lambda: (np.fix(a / b).astype(np.int64) if real is CType.Int else a / b),
             "mod": lambda: np.fmod(a, b), "neg": lambda: -a,
             "inv": lambda: np.asarray(1, dtype=b.dtype) / b,
             "max": lambda: np.where(a < b, b, a), "min": lambda: np.where(b < a, b, a),
             "abs": lambda: np.abs(a), "select": lambda: np.where(a < b, a, b),
             "eq": lambda: a == b, "ne": lambda: a != b, "lt": lambda: a < b,
             "le": lambda: a <= b, "gt": lambda: a > b, "ge": lambda: a >= b,
             "and": lambda: (a < 0) & (b > 4), "or": lambda: (a < 0) | (b > 4),
             "not": lambda: ~(a < 0), "sqrt": lambda: np.sqrt(b),
             "cast_int": lambda: (b * np.asarray(100, dtype=b.dtype)).astype(np.int32),
             "cast_other": lambda: a.astype(dt)}[op]()
    return np.asarray(r).astype(dt)
