"""Lowered SSA form of one OM kernel for the HIP backend.

The OM graph (``OM/Graph.hs``) is bipartite and, because the Builder never
shares sub-expressions, full of duplicates.  The HIP backend first turns each
kernel into a compact SSA list of value operations with:

* value numbering (identical operation on identical arguments -> one value),
* exact constant folding (integers, booleans, and IEEE float operations done
  in the value's own precision with numpy, so folded bits equal run-time bits),
* a symbolic range analysis for ``LoadIndex``/``LoadSize`` on Cyclic axes:
  an index lies in ``[0, n-1]`` and the size is ``n``, which decides
  comparisons such as the out-of-domain test of a boundary condition.

Folding never changes results: every rewrite is value-preserving for all grid
sizes (IEEE operations are deterministic; comparisons are decided only when
they hold for every ``n >= 1``).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

from ... import annotation as A
from ...om.arith import COMPARISONS, Op, ReduceOp
from ...om.dyn import CType, Realm
from ...om.graph import (Arith, Broadcast, Graph, Imm, Load, LoadIndex, LoadSize,
                         Reduce, Shift, Store)


@dataclass
class LNode:
    id: int
    op: str                    # load imm index size shift arith reduce bcast
    args: Tuple[int, ...]
    realm: Realm
    ctype: CType
    payload: object = None     # static idx / literal / axis / vec / (Op, cast) / ReduceOp
    hint_manifest: bool = False
    om_nodes: List[int] = field(default_factory=list)

    @property
    def arith_op(self) -> Optional[Op]:
        return self.payload[0] if self.op == "arith" else None


@dataclass
class LKernel:
    name: str
    nodes: List[LNode]
    stores: List[Tuple[int, int]]           # (static idx, value id)
    ndim: int

    def users(self) -> List[List[int]]:
        u = [[] for _ in self.nodes]
        for n in self.nodes:
            for a in n.args:
                u[a].append(n.id)
        return u


# ------------------------------------------------------------ range analysis
class _Rng:
    """Interval [lo, hi] with bounds k*n_axis + c (axis None: plain constants)."""
    __slots__ = ("axis", "lo", "hi")

    def __init__(self, axis, lo, hi):
        self.axis, self.lo, self.hi = axis, lo, hi


def _lt_always(a: _Rng, b: _Rng) -> Optional[bool]:
    """True if a < b for every n >= 1, False if a >= b for every n >= 1, else None."""
    if a.axis is not None and b.axis is not None and a.axis != b.axis:
        return None
    dk, dc = a.hi[0] - b.lo[0], a.hi[1] - b.lo[1]
    if dk <= 0 and dk + dc < 0:
        return True
    dk, dc = a.lo[0] - b.hi[0], a.lo[1] - b.hi[1]
    if dk >= 0 and dk + dc >= 0:
        return False
    return None


def _np_scalar(v, ct: CType):
    if ct is CType.Float:
        return np.float32(v)
    if ct is CType.Double:
        return np.float64(v)
    if ct is CType.Int:
        return np.int32(v)
    return bool(v)


def _fold(op: Op, vals, ct_args: List[CType], ct_out: CType, cast_to):
    """Exact evaluation of a pure operation on constants, or None."""
    with np.errstate(all="ignore"):
        a = [_np_scalar(v, c) for v, c in zip(vals, ct_args)]
        try:
            if op is Op.Add:
                r = a[0] + a[1]
            elif op is Op.Sub:
                r = a[0] - a[1]
            elif op is Op.Mul:
                r = a[0] * a[1]
            elif op is Op.Div:
                if ct_args[0] is CType.Int:
                    if int(a[1]) == 0:
                        return None
                    q = abs(int(a[0])) // abs(int(a[1]))
                    r = np.int32(q if (int(a[0]) >= 0) == (int(a[1]) >= 0) else -q)
                else:
                    r = a[0] / a[1]
            elif op is Op.Mod:
                if ct_args[0] is not CType.Int or int(a[1]) == 0:
                    return None
                q = abs(int(a[0])) % abs(int(a[1]))
                r = np.int32(q if int(a[0]) >= 0 else -q)
            elif op is Op.Neg:
                r = -a[0]
            elif op is Op.Inv:
                if not ct_args[0].is_float:
                    return None
                r = _np_scalar(1, ct_args[0]) / a[0]
            elif op is Op.Not:
                r = not a[0]
            elif op is Op.And:
                r = bool(a[0]) and bool(a[1])
            elif op is Op.Or:
                r = bool(a[0]) or bool(a[1])
            elif op in COMPARISONS:
                r = {Op.EQ: a[0] == a[1], Op.NE: a[0] != a[1], Op.LT: a[0] < a[1],
                     Op.LE: a[0] <= a[1], Op.GT: a[0] > a[1], Op.GE: a[0] >= a[1]}[op]
                r = bool(r)
            elif op is Op.Max:
                r = a[1] if a[0] < a[1] else a[0]          # std::max
            elif op is Op.Min:
                r = a[1] if a[1] < a[0] else a[0]          # std::min
            elif op is Op.Select:
                r = a[1] if a[0] else a[2]
            elif op is Op.Sqrt and ct_args[0].is_float:
                r = np.sqrt(a[0])
            elif op is Op.Cast:
                if cast_to is CType.Int:
                    r = np.int32(int(a[0]))
                elif cast_to is CType.Bool:
                    r = bool(a[0])
                else:
                    r = _np_scalar(a[0], cast_to)
            else:
                return None
        except (OverflowError, ValueError, ZeroDivisionError):
            return None
    return _np_scalar(r, ct_out) if ct_out is not CType.Bool else bool(r)


# ------------------------------------------------------------ building
def lower_kernel(om, ki: int, cyclic: Tuple[bool, ...]) -> LKernel:
    g: Graph = om.kernels[ki].graph
    nodes: List[LNode] = []
    memo: Dict[tuple, int] = {}
    val_map: Dict[int, int] = {}           # OM value id -> LNode id
    const: Dict[int, object] = {}
    rng: Dict[int, _Rng] = {}

    def emit(op, args, realm, ct, payload, om_id, hint=False) -> int:
        key = (op, args, realm, ct, _hashable(payload))
        if key in memo:
            nid = memo[key]
            nodes[nid].om_nodes.append(om_id)
            nodes[nid].hint_manifest |= hint
            return nid
        nid = len(nodes)
        nodes.append(LNode(nid, op, args, realm, ct, payload, hint, [om_id]))
        memo[key] = nid
        return nid

    def make_const(value, realm, ct, om_id) -> int:
        v = _np_scalar(value, ct) if ct is not CType.Bool else bool(value)
        nid = emit("imm", (), realm, ct, v, om_id)
        const[nid] = v
        if ct is not CType.Bool:
            rng[nid] = _Rng(None, (0, v), (0, v))
        return nid

    stores = []
    for v, node in enumerate(g.nodes):
        if not node.is_value:
            if isinstance(node.inst, Store):
                stores.append((node.inst.static, val_map[g.pre(v)[0]]))
            continue
        _, inst = g.inst_of_value(v)
        inst_id = g.pre(v)[0]
        args = tuple(val_map[p] for p in g.pre(inst_id))
        realm, ct = node.dyn.realm, node.dyn.ctype
        hint = A.to_maybe(A.Allocation, node.anot) is A.Allocation.Manifest
        if isinstance(inst, Load):
            nid = emit("load", (), realm, ct, inst.static, v)
        elif isinstance(inst, Imm):
            nid = make_const(inst.value, realm, ct, v)
        elif isinstance(inst, LoadIndex):
            nid = emit("index", (), realm, ct, inst.axis, v)
            if cyclic[inst.axis]:
                rng[nid] = _Rng(inst.axis, (0, 0), (1, -1))
        elif isinstance(inst, LoadSize):
            nid = emit("size", (), realm, ct, inst.axis, v)
            rng[nid] = _Rng(inst.axis, (1, 0), (1, 0))
        elif isinstance(inst, Shift):
            if args[0] in const:
                nid = make_const(const[args[0]], realm, ct, v)
            elif all(x == 0 for x in inst.vec):
                nid = args[0]
                nodes[nid].om_nodes.append(v)
            else:
                nid = emit("shift", args, realm, ct, inst.vec, v)
        elif isinstance(inst, Broadcast):
            if args[0] in const:
                nid = make_const(const[args[0]], realm, ct, v)
            else:
                nid = emit("bcast", args, realm, ct, None, v)
                if args[0] in rng:
                    rng[nid] = rng[args[0]]
        elif isinstance(inst, Reduce):
            nid = emit("reduce", args, realm, ct, inst.op, v)
        elif isinstance(inst, Arith):
            op = inst.op
            nid = None
            if all(a in const for a in args):
                r = _fold(op, [const[a] for a in args], [nodes[a].ctype for a in args], ct,
                          inst.cast_to)
                if r is not None:
                    nid = make_const(r, realm, ct, v)
            if nid is None and op is Op.Select and args[0] in const:
                nid = args[1] if const[args[0]] else args[2]
                nodes[nid].om_nodes.append(v)
            if nid is None and op in (Op.And, Op.Or):
                for a, b in ((args[0], args[1]), (args[1], args[0])):
                    if a in const:
                        if (op is Op.And and not const[a]) or (op is Op.Or and const[a]):
                            nid = make_const(const[a], realm, ct, v)
                        else:
                            nid = b
                            nodes[nid].om_nodes.append(v)
                        break
            if nid is None and op in COMPARISONS and args[0] in rng and args[1] in rng:
                ra, rb = rng[args[0]], rng[args[1]]
                dec = None
                if op is Op.LT:
                    dec = _lt_always(ra, rb)
                elif op is Op.GE:
                    d = _lt_always(ra, rb)
                    dec = None if d is None else (not d)
                elif op is Op.GT:
                    dec = _lt_always(rb, ra)
                elif op is Op.LE:
                    d = _lt_always(rb, ra)
                    dec = None if d is None else (not d)
                if dec is not None:
                    nid = make_const(dec, realm, ct, v)
            if nid is None:
                nid = emit("arith", args, realm, ct, (op, inst.cast_to), v, hint)
                if op is Op.Cast and args[0] in rng and inst.cast_to is not CType.Bool:
                    rng[nid] = rng[args[0]]
        else:
            raise TypeError(inst)
        val_map[v] = nid
    return _prune(LKernel(om.kernels[ki].name, nodes, stores, om.ndim))


def _hashable(p):
    if isinstance(p, np.generic):
        return (type(p).__name__, p.tobytes())
    return p


def _prune(k: LKernel) -> LKernel:
    """Drop values no store reaches; renumber densely (topological order kept)."""
    live = set()
    stack = [v for _, v in k.stores]
    while stack:
        x = stack.pop()
        if x in live:
            continue
        live.add(x)
        stack.extend(k.nodes[x].args)
    remap = {}
    out = []
    for n in k.nodes:
        if n.id in live:
            remap[n.id] = len(out)
            out.append(LNode(len(out), n.op, tuple(remap[a] for a in n.args), n.realm, n.ctype,
                             n.payload, n.hint_manifest, n.om_nodes))
    return LKernel(k.name, out, [(s, remap[v]) for s, v in k.stores], k.ndim)
