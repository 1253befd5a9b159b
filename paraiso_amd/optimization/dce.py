"""Dead-code elimination (``Optimization/DeadCodeElimination.hs:19-70``).

A node is alive if it is a Store or a Load instruction, or if any successor
is alive.  Dead nodes are removed and the survivors renumbered densely in
their original order.
"""
from __future__ import annotations

from .. import annotation as A
from ..om.graph import Graph, Load, Store


def dead_code_elimination(g: Graph) -> Graph:
    n = len(g)
    alive = [False] * n
    # successors always have larger ids than their predecessors
    for i in range(n - 1, -1, -1):
        inst = g.nodes[i].inst
        if isinstance(inst, (Store, Load)):
            alive[i] = True
        else:
            alive[i] = any(alive[j] for j in g.succ(i))
    old_to_new = {}
    out = Graph()
    for i in range(n):
        if alive[i]:
            old_to_new[i] = len(old_to_new)
    for i in range(n):
        if not alive[i]:
            continue
        node = g.nodes[i].copy()
        node.anot = A.set_(A.Alive(True), node.anot)
        preds = [old_to_new[p] for p in g.preds[i] if p in old_to_new]
        out.add(preds, node)
    return out
