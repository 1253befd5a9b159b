"""Allocation decision (``Optimization/DecideAllocation.hs:25-83``).

Values right after a Load are Existing; values right before a Store, before
or after a Reduce, before or after a Broadcast are Manifest; everything else
is Delayed unless the user already annotated it (``weakSet``), and value
nodes record the choice ``[Delayed, Manifest]`` for tuners.
"""
from __future__ import annotations

from .. import annotation as A
from ..om.graph import Broadcast, Graph, Load, Reduce, Store


def decide_allocation(g: Graph) -> Graph:
    out = g.copy()
    for i, node in enumerate(g.nodes):
        preds = g.pre(i)
        pre0 = g.nodes[preds[0]] if preds else None
        sucs = [g.nodes[j] for j in g.succ(i)]
        pre_inst = pre0.inst if pre0 is not None else None

        after_load = isinstance(pre_inst, Load)
        before_store = any(isinstance(s.inst, Store) for s in sucs)
        before_reduce = any(isinstance(s.inst, Reduce) for s in sucs)
        after_reduce = isinstance(pre_inst, Reduce)
        before_bcast = any(isinstance(s.inst, Broadcast) for s in sucs)
        after_bcast = isinstance(pre_inst, Broadcast)

        anot = out.nodes[i].anot
        if after_load:
            anot = A.set_(A.Allocation.Existing, anot)
        elif before_store or before_reduce or after_reduce or before_bcast or after_bcast:
            anot = A.set_(A.Allocation.Manifest, anot)
        else:
            if node.is_value:
                anot = A.set_(A.AllocationChoice((A.Allocation.Delayed, A.Allocation.Manifest)), anot)
            anot = A.weak_set(A.Allocation.Delayed, anot)
        out.nodes[i].anot = anot
    return out


def allocation_of(node) -> A.Allocation:
    a = A.to_maybe(A.Allocation, node.anot)
    if a is None:
        raise ValueError("allocation decided before use")
    return a
