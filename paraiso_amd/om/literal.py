"""Printing immediates as C/HIP literals.

Follows the literal database of ``Language/Paraiso/Generator/ClarisTrans.hs:180-199``:
``Double`` prints with Haskell ``show`` (shortest round-trip digits, fixed
notation for 0.1 <= |x| < 10^7, otherwise ``d.ddde<exp>``), ``Float`` prints the
same way with an ``f`` suffix, ``Int`` in decimal, ``Bool`` as ``true``/``false``.
Any shortest round-trip decimal parses back to the same binary value, so the
printed literal carries exactly the bits of the immediate.
"""
from __future__ import annotations

import math

import numpy as np

from .dyn import CType


def _digits(x, ctype: CType):
    """Shortest round-trip digits ``ds`` and exponent ``e`` with x = 0.ds * 10^e."""
    v = np.float32(x) if ctype is CType.Float else np.float64(x)
    if v == 0:
        return "0", 0
    s = np.format_float_scientific(abs(v), unique=True, trim="-")
    mant, exp = s.split("e")
    ds = mant.replace(".", "")
    ds = ds.rstrip("0") or "0"
    return ds, int(exp) + 1


def haskell_show_float(x, ctype: CType) -> str:
    v = float(x)
    if math.isnan(v):
        return "NaN"
    if math.isinf(v):
        return "Infinity" if v > 0 else "-Infinity"
    neg = math.copysign(1.0, v) < 0 and v != 0
    ds, e = _digits(abs(v), ctype)
    if e < 0 or e > 7:
        if len(ds) == 1:
            body = ds + ".0e" + str(e - 1)
        else:
            body = ds[0] + "." + ds[1:] + "e" + str(e - 1)
    elif e <= 0:
        body = "0." + "0" * (-e) + ds
    else:
        if len(ds) <= e:
            body = ds + "0" * (e - len(ds)) + ".0"
        else:
            body = ds[:e] + "." + ds[e:]
    return ("-" if neg else "") + body


def c_literal(value, ctype: CType) -> str:
    if ctype is CType.Bool:
        return "true" if value else "false"
    if ctype is CType.Int:
        return str(int(value))
    if ctype is CType.Float:
        s = haskell_show_float(value, ctype)
        if s in ("NaN", "Infinity", "-Infinity"):
            return {"NaN": "(0.0f/0.0f)", "Infinity": "(1.0f/0.0f)",
                    "-Infinity": "(-1.0f/0.0f)"}[s]
        return s + "f"
    s = haskell_show_float(value, ctype)
    if s in ("NaN", "Infinity", "-Infinity"):
        return {"NaN": "(0.0/0.0)", "Infinity": "(1.0/0.0)", "-Infinity": "(-1.0/0.0)"}[s]
    return s
