"""Dependency analysis and write grouping
(``Optimization/DependencyAnalysis.hs:37-258``).

Strict nodes are the Manifest and Existing ones.  For every node:
* Direct   = strict nodes read by the subroutine computing it,
* Indirect = strict nodes that must be finished before it,
* Calc     = all nodes evaluated by the subroutine computing it.
Manifest nodes are grouped greedily in id order: a node joins the first
existing group all of whose members it can coexist with (no indirect
dependency, same realm, same Valid region), else opens group
``1 + max(groups of earlier Manifest nodes)``.  Group numbers are then made
unique across the kernels of the machine (``OMWriteGroup``).
"""
from __future__ import annotations

from .. import annotation as A
from ..om.graph import OM, Graph, Kernel
from .allocation import allocation_of
from .boundary import valid_of


def _bits(mask: int):
    out = []
    k = 0
    while mask:
        if mask & 1:
            out.append(k)
        mask >>= 1
        k += 1
    return out


def dependency_analysis(g: Graph) -> Graph:
    n = len(g)
    alloc = [allocation_of(nd) for nd in g.nodes]
    strict = [a in (A.Allocation.Manifest, A.Allocation.Existing) for a in alloc]
    sidx_to_idx = [i for i in range(n) if strict[i]]
    idx_to_sidx = {i: s for s, i in enumerate(sidx_to_idx)}

    dep_w = [0] * n
    dep_r = [0] * n
    ind_w = [0] * n
    ind_r = [0] * n
    calc_w = [frozenset()] * n
    calc_r = [frozenset()] * n
    for i in range(n):
        dw = 0
        iw = 0
        cw = set()
        for p in g.pre(i):
            dw |= dep_r[p]
            iw |= ind_r[p]
            cw |= calc_r[p]
        dep_w[i] = dw
        ind_w[i] = iw
        calc_w[i] = frozenset(cw)
        if strict[i]:
            me = 1 << idx_to_sidx[i]
            dep_r[i] = me
            ind_r[i] = me | iw
            calc_r[i] = frozenset([i])
        else:
            dep_r[i] = dw
            ind_r[i] = iw
            calc_r[i] = frozenset([i]) | calc_w[i]

    def to_nodes(mask):
        return tuple(sidx_to_idx[s] for s in _bits(mask))

    realm = [nd.dyn.realm if nd.is_value else None for nd in g.nodes]
    valid = [valid_of(nd) for nd in g.nodes]
    manifest_nodes = [i for i in range(n) if alloc[i] is A.Allocation.Manifest]

    def coexist(i, j):
        if i == j:
            return True
        if i < j:
            i, j = j, i
        dependent = bool(ind_w[i] >> idx_to_sidx[j] & 1)
        same = realm[i] == realm[j] and valid[i] == valid[j]
        return (not dependent) and same

    group = {}
    for idx in manifest_nodes:
        pres = [m for m in manifest_nodes if m < idx]
        if not pres:
            group[idx] = 0
            continue
        existing = sorted(set(group[m] for m in pres))
        co = [gid for gid in existing
              if all(coexist(idx, m) for m in pres if group[m] == gid)]
        group[idx] = co[0] if co else 1 + max(group[m] for m in pres)

    out = g.copy()
    for i in range(n):
        anot = out.nodes[i].anot
        anot = A.set_(A.Indirect(to_nodes(ind_w[i])), anot)
        anot = A.set_(A.Direct(to_nodes(dep_w[i])), anot)
        anot = A.set_(A.Calc(calc_w[i]), anot)
        if alloc[i] is A.Allocation.Manifest:
            anot = A.set_(A.KernelWriteGroup(group[i]), anot)
        out.nodes[i].anot = anot
    return out


def write_grouping(om: OM) -> OM:
    graphs = [dependency_analysis(k.graph) for k in om.kernels]
    counts = []
    for g in graphs:
        ids = [A.to_maybe(A.KernelWriteGroup, nd.anot) for nd in g.nodes]
        ids = [x.gid for x in ids if x is not None]
        counts.append(1 + max([-1] + ids))
    kernels = []
    for ki, (k, g) in enumerate(zip(om.kernels, graphs)):
        diff = sum(counts[:ki])
        for nd in g.nodes:
            nd.anot = [A.OMWriteGroup(a.gid + diff) if isinstance(a, A.KernelWriteGroup) else a
                       for a in nd.anot]
        kernels.append(Kernel(k.name, g))
    return OM(om.name, om.setup, kernels, om.ndim)
