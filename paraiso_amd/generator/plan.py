"""The Plan: storages, subkernels and margins of an optimised OM.

Restates ``Language/Paraiso/Generator/OMTrans.hs:39-206`` and
``Generator/Plan.hs:33-139``:
* storages = the statics followed by every Manifest node (kernel order, id order),
* one subkernel per OM write group, with its inputs (union of Direct), calcs
  (union of Calc), outputs, realm and lower/upper boundary,
* the plan margins are the intersection of the Valid regions of all Store
  instructions, forced to 0 on Cyclic axes.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Tuple

from .. import annotation as A
from ..om.dyn import DynValue, Realm
from ..om.graph import OM, Store
from ..optimization import optimize
from .native import Setup


@dataclass(frozen=True)
class StorageIdx:
    kind: str            # "static" | "manifest"
    a: int               # static index, or kernel index
    b: int = -1          # node index for manifest storages


@dataclass
class StorageRef:
    idx: StorageIdx
    dyn: DynValue
    name: str


@dataclass
class SubKernel:
    kernel_idx: int
    group: int
    inputs: List[int]
    calcs: List[int]
    outputs: List[int]
    realm: Realm
    lower_boundary: Tuple[int, ...]
    upper_boundary: Tuple[int, ...]
    kernel_name: str

    @property
    def name(self) -> str:
        return "om_%s_sub_%d" % (self.kernel_name, self.group)


@dataclass
class Plan:
    name: str
    om: OM
    storages: List[StorageRef]
    subkernels: List[SubKernel]
    lower_margin: Tuple[int, ...]
    upper_margin: Tuple[int, ...]
    setup: Setup

    def memory_size(self) -> Tuple[int, ...]:
        return tuple(n + l + u for n, l, u in zip(self.setup.local_size,
                                                 self.lower_margin, self.upper_margin))

    def storage_of_static(self, sidx: int) -> StorageRef:
        return next(s for s in self.storages if s.idx == StorageIdx("static", sidx))

    def storage_of_manifest(self, kidx: int, node: int) -> StorageRef:
        return next(s for s in self.storages if s.idx == StorageIdx("manifest", kidx, node))


def _valid_to_lower(valid: A.Valid, setup: Setup) -> Tuple[int, ...]:
    out = []
    for ax, iv in enumerate(valid.intervals):
        if setup.boundary[ax] is A.Condition.Cyclic:
            out.append(0)
            continue
        lo = iv.lower
        if lo is None:
            raise ValueError("wrong lower Margin! (empty interval)")
        if lo.tag == A.NearBoundary.NEG:
            out.append(0)
        elif lo.tag == A.NearBoundary.LOW:
            out.append(lo.x)
        else:
            raise ValueError("wrong lower Margin!")
    return tuple(out)


def _valid_to_upper(valid: A.Valid, setup: Setup) -> Tuple[int, ...]:
    out = []
    for ax, iv in enumerate(valid.intervals):
        if setup.boundary[ax] is A.Condition.Cyclic:
            out.append(0)
            continue
        up = iv.upper
        if up is None:
            raise ValueError("wrong upper Margin! (empty interval)")
        if up.tag == A.NearBoundary.POS:
            out.append(0)
        elif up.tag == A.NearBoundary.UP:
            out.append(-up.x)
        else:
            raise ValueError("wrong upper Margin!")
    return tuple(out)


def translate(setup: Setup, om0: OM) -> Plan:
    """``OMTrans.translate``: optimise at the setup's level, then plan."""
    if om0.ndim != setup.ndim:
        raise ValueError("OM has %d dimensions, setup has %d" % (om0.ndim, setup.ndim))
    om = optimize(setup.opt_level, om0)
    storages = [StorageRef(StorageIdx("static", i), s.dyn, "om_s%d_%s" % (i, s.name))
                for i, s in enumerate(om.setup.static_values)]
    manifest = []          # (kernel idx, node idx)
    for ki, k in enumerate(om.kernels):
        for i, nd in enumerate(k.graph.nodes):
            if A.to_maybe(A.Allocation, nd.anot) is A.Allocation.Manifest:
                if not nd.is_value:
                    raise ValueError("a non-Value node is marked as Manifest")
                manifest.append((ki, i))
                storages.append(StorageRef(StorageIdx("manifest", ki, i), nd.dyn,
                                           "om_m%d_%d" % (ki, i)))

    def group_of(ki, i):
        g = A.to_maybe(A.OMWriteGroup, om.kernels[ki].graph.nodes[i].anot)
        if g is None:
            raise ValueError("OMWriteGroup missing: %d %d" % (ki, i))
        return g.gid

    groups = {}
    for ki, i in manifest:
        groups.setdefault(group_of(ki, i), []).append((ki, i))
    n_sub = 1 + max([-1] + list(groups))
    subs = []
    for gid in range(n_sub):
        members = groups[gid]
        ki = members[0][0]
        g = om.kernels[ki].graph
        ins, calcs = set(), set()
        realms, valids = [], []
        for _, i in members:
            nd = g.nodes[i]
            d = A.to_maybe(A.Direct, nd.anot)
            c = A.to_maybe(A.Calc, nd.anot)
            if d is not None:
                ins |= set(d.nodes)
            if c is not None:
                calcs |= set(c.nodes)
            realms.append(nd.dyn.realm)
            valids.extend(A.to_list(A.Valid, nd.anot))
        if any(r != realms[0] for r in realms) or any(v != valids[0] for v in valids):
            raise ValueError("elements are different. mismatch.")
        subs.append(SubKernel(ki, gid, sorted(ins), sorted(calcs), [i for _, i in members],
                              realms[0], _valid_to_lower(valids[0], setup),
                              _valid_to_upper(valids[0], setup), om.kernels[ki].name))

    store_valids = []
    for k in om.kernels:
        for nd in k.graph.nodes:
            if isinstance(nd.inst, Store):
                store_valids.extend(A.to_list(A.Valid, nd.anot))
    if not store_valids:
        raise ValueError("the OM stores nothing")
    united = store_valids[0]
    for v in store_valids[1:]:
        united = united.intersection(v)
    return Plan(om.name, om, storages, subs, _valid_to_lower(united, setup),
                _valid_to_upper(united, setup), setup)
