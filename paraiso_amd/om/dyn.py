"""Realms, content types and the type unifier of the Orthotope Machine.

Follows ``Language/Paraiso/OM/Realm.hs:37-40`` (``Scalar | Array``) and
``Language/Paraiso/OM/DynValue.hs:17`` (``DynValue {realm, typeRep}``).

The Haskell builder gets the realm and content type of every node from GHC's
type inference: an immediate such as ``1`` in ``gen + 1`` takes the type of
``gen`` (``OM/Builder/Internal.hs:132-142,274-277``).  Python has no such
inference, so every builder value carries a pair of union-find variables that
are unified as operations combine values; a kernel is finished only when every
node's pair is resolved.
"""
from __future__ import annotations

from dataclasses import dataclass
from enum import Enum
from fractions import Fraction

import numpy as np


class Realm(Enum):
    Scalar = "Scalar"
    Array = "Array"

    def old_name(self) -> str:
        # the older OM dumps call the realms Global / Local
        return "Global" if self is Realm.Scalar else "Local"


class CType(Enum):
    """Content types the generators know (``ClarisTrans.hs:180-199``)."""
    Bool = "Bool"
    Int = "Int"
    Float = "Float"
    Double = "Double"

    @property
    def c_name(self) -> str:
        return {"Bool": "bool", "Int": "int", "Float": "float",
                "Double": "double"}[self.value]

    @property
    def np_dtype(self):
        return {"Bool": np.uint8, "Int": np.int32, "Float": np.float32,
                "Double": np.float64}[self.value]

    @property
    def size(self) -> int:
        return np.dtype(self.np_dtype).itemsize

    @property
    def is_float(self) -> bool:
        return self in (CType.Float, CType.Double)


@dataclass(frozen=True)
class DynValue:
    realm: Realm
    ctype: CType

    def show(self, old: bool = False) -> str:
        r = self.realm.old_name() if old else self.realm.value
        return "DynValue {realm = %s, typeRep = %s}" % (r, self.ctype.value)


class TyVar:
    """Union-find cell holding either a concrete value or nothing."""
    __slots__ = ("parent", "value")

    def __init__(self, value=None):
        self.parent = None
        self.value = value

    def find(self) -> "TyVar":
        r = self
        while r.parent is not None:
            r = r.parent
        # path compression
        s = self
        while s.parent is not None and s.parent is not r:
            nxt = s.parent
            s.parent = r
            s = nxt
        return r

    def get(self):
        return self.find().value

    def unify(self, other: "TyVar", what: str = "type") -> None:
        a, b = self.find(), other.find()
        if a is b:
            return
        if a.value is not None and b.value is not None and a.value != b.value:
            raise TypeError("OM type mismatch in %s: %s vs %s" % (what, a.value, b.value))
        if a.value is None:
            a.parent = b
        else:
            b.parent = a


def fraction_to_float32(fr: Fraction) -> np.float32:
    """Correctly rounded rational -> binary32 (Haskell ``fromRational :: Float``).

    Rounding through binary64 first can double-round; the neighbours of the
    binary64 guess are compared exactly instead (ties to even).
    """
    guess = np.float32(float(fr))
    best = guess
    cands = [np.nextafter(guess, np.float32(-np.inf)), guess,
             np.nextafter(guess, np.float32(np.inf))]

    def dist(c):
        return abs(Fraction(float(c)) - fr)

    for c in cands:
        if not np.isfinite(c):
            continue
        dc, db = dist(c), dist(best)
        if dc < db or (dc == db and (int(c.view(np.uint32)) & 1) == 0 and c != best):
            best = c
    return np.float32(best)


def fraction_to_float64(fr: Fraction) -> float:
    # Fraction.__float__ is correctly rounded (numerator/denominator true division)
    return float(fr)


def convert_literal(value, ctype: CType):
    """Turn a Python literal (int, float, Fraction, bool) into a typed value."""
    if ctype is CType.Bool:
        return bool(value)
    if ctype is CType.Int:
        if isinstance(value, float) and not float(value).is_integer():
            raise TypeError("non-integral literal %r used as Int" % (value,))
        return int(value)
    fr = value if isinstance(value, Fraction) else Fraction(value)
    if ctype is CType.Float:
        return fraction_to_float32(fr)
    return np.float64(fraction_to_float64(fr))
