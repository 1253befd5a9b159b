"""Arithmetic and reduction operators of the OM.

The operator set and arities follow ``Language/Paraiso/OM/Arithmetic.hs:20-110``;
reduction operators follow ``Language/Paraiso/OM/Reduce.hs:9-16``.
"""
from __future__ import annotations

from enum import Enum


class Op(Enum):
    Identity = "Identity"
    Add = "Add"
    Sub = "Sub"
    Neg = "Neg"
    Mul = "Mul"
    Div = "Div"
    Mod = "Mod"
    DivMod = "DivMod"
    Inv = "Inv"
    Not = "Not"
    And = "And"
    Or = "Or"
    EQ = "EQ"
    NE = "NE"
    LT = "LT"
    LE = "LE"
    GT = "GT"
    GE = "GE"
    Max = "Max"
    Min = "Min"
    Abs = "Abs"
    Signum = "Signum"
    Select = "Select"
    Ipow = "Ipow"
    Pow = "Pow"
    Madd = "Madd"
    Msub = "Msub"
    Nmadd = "Nmadd"
    Nmsub = "Nmsub"
    Sqrt = "Sqrt"
    Exp = "Exp"
    Log = "Log"
    Sin = "Sin"
    Cos = "Cos"
    Tan = "Tan"
    Asin = "Asin"
    Acos = "Acos"
    Atan = "Atan"
    Atan2 = "Atan2"
    Sincos = "Sincos"
    Cast = "Cast"


# (inputs, outputs) per Arithmetic.hs:73-110
ARITY = {
    Op.Identity: (1, 1), Op.Add: (2, 1), Op.Sub: (2, 1), Op.Neg: (1, 1),
    Op.Mul: (2, 1), Op.Div: (2, 1), Op.Mod: (2, 1), Op.DivMod: (2, 2),
    Op.Inv: (1, 1), Op.Not: (1, 1), Op.And: (2, 1), Op.Or: (2, 1),
    Op.EQ: (2, 1), Op.NE: (2, 1), Op.LT: (2, 1), Op.LE: (2, 1),
    Op.GT: (2, 1), Op.GE: (2, 1), Op.Max: (2, 1), Op.Min: (2, 1),
    Op.Abs: (1, 1), Op.Signum: (1, 1), Op.Select: (3, 1), Op.Ipow: (2, 1),
    Op.Pow: (2, 1), Op.Madd: (3, 1), Op.Msub: (3, 1), Op.Nmadd: (3, 1),
    Op.Nmsub: (3, 1), Op.Sqrt: (1, 1), Op.Exp: (1, 1), Op.Log: (1, 1),
    Op.Sin: (1, 1), Op.Cos: (1, 1), Op.Tan: (1, 1), Op.Asin: (1, 1),
    Op.Acos: (1, 1), Op.Atan: (1, 1), Op.Atan2: (2, 1), Op.Sincos: (1, 2),
    Op.Cast: (1, 1),
}

COMPARISONS = (Op.EQ, Op.NE, Op.LT, Op.LE, Op.GT, Op.GE)


class ReduceOp(Enum):
    Max = "Max"
    Min = "Min"
    Sum = "Sum"

    def to_arith(self) -> Op:
        return {ReduceOp.Max: Op.Max, ReduceOp.Min: Op.Min,
                ReduceOp.Sum: Op.Add}[self]
