"""Valid-region analysis (``Optimization/BoundaryAnalysis.hs:33-103``).

Load -> full array; Shift v -> full intersected with the argument's region
moved by v; Arith -> intersection of the arguments; Imm, Reduce, Broadcast,
LoadIndex, LoadSize and Scalar values -> infinite; Store and Array values ->
their single predecessor's region.
"""
from __future__ import annotations

from .. import annotation as A
from ..om.dyn import Realm
from ..om.graph import (Arith, Broadcast, Graph, Imm, Load, LoadIndex, LoadSize,
                        Reduce, Shift, Store)


def boundary_analysis(g: Graph, ndim: int) -> Graph:
    full = A.full_valid(ndim)
    inf = A.infinite_valid(ndim)
    memo = [None] * len(g)
    for i, node in enumerate(g.nodes):
        preds = g.pre(i)
        if node.is_value:
            if node.dyn.realm is Realm.Scalar:
                v = inf
            else:
                if len(preds) != 1:
                    raise ValueError("node[%d] only 1 pre expected" % i)
                v = memo[preds[0]]
        else:
            inst = node.inst
            if isinstance(inst, (Imm, Reduce, Broadcast, LoadIndex, LoadSize)):
                v = inf
            elif isinstance(inst, Load):
                v = full
            elif isinstance(inst, Store):
                v = memo[preds[0]]
            elif isinstance(inst, Shift):
                pre = memo[preds[0]]
                moved = A.Valid(tuple(
                    A.EMPTY if iv.empty else A.Interval(iv.lower.shifted(d), iv.upper.shifted(d))
                    for d, iv in zip(inst.vec, pre.intervals)))
                v = full.intersection(moved)
            elif isinstance(inst, Arith):
                if not preds:
                    raise ValueError("arith node[%d] has 0 pre" % i)
                v = memo[preds[0]]
                for p in preds[1:]:
                    v = v.intersection(memo[p])
            else:
                raise TypeError(inst)
        memo[i] = v
    out = g.copy()
    for i in range(len(g)):
        out.nodes[i].anot = A.set_(memo[i], out.nodes[i].anot)
    return out


def valid_of(node) -> A.Valid:
    v = A.to_maybe(A.Valid, node.anot)
    if v is None:
        raise ValueError("boundaryAnalysis must run first")
    return v
