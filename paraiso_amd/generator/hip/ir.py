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

With ``exact=True`` (the fast FP32 builds) floating constants are instead
folded in exact rational arithmetic from the literals the program wrote and
rounded once: ``1.4 - 1`` becomes the binary32 nearest 2/5 rather than the
difference of the rounded 1.4 and 1, and ``1 / (1.4 - 1)`` becomes exactly 2.5.
The exact value of each floating constant is kept (``LKernel.exact``) so the
emitter can multiply by it with a compensated product.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from fractions import Fraction
from typing import Dict, List, Optional, Tuple

import numpy as np

from ... import annotation as A
from ...om.arith import COMPARISONS, Op
from ...om.dyn import CType, Realm, fraction_to_float32, fraction_to_float64
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
    # exact rational value of each floating constant (exact-folding builds only)
    exact: Dict[int, Fraction] = field(default_factory=dict)

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


def round_exact(x: Fraction, ct: CType):
    """The value of type ``ct`` nearest the rational ``x`` (ties to even)."""
    return fraction_to_float32(x) if ct is CType.Float else np.float64(fraction_to_float64(x))


def _fold_exact(op: Op, xs: List[Fraction], ct_args: List[CType], cast_to) -> Optional[Fraction]:
    """Exact rational value of an operation on exact constants, or None."""
    try:
        if op is Op.Add:
            return xs[0] + xs[1]
        if op is Op.Sub:
            return xs[0] - xs[1]
        if op is Op.Mul:
            return xs[0] * xs[1]
        if op is Op.Div:
            return xs[0] / xs[1]
        if op is Op.Neg:
            return -xs[0]
        if op is Op.Inv:
            return 1 / xs[0]
        if op is Op.Cast and cast_to is not None and cast_to.is_float:
            return xs[0]
    except ZeroDivisionError:
        return None
    return None


# ------------------------------------------------------------ building
def lower_kernel(om, ki: int, cyclic: Tuple[bool, ...], exact: bool = False) -> LKernel:
    """``exact``: fold floating constants in exact rational arithmetic (see above)."""
    g: Graph = om.kernels[ki].graph
    nodes: List[LNode] = []
    memo: Dict[tuple, int] = {}
    val_map: Dict[int, int] = {}           # OM value id -> LNode id
    const: Dict[int, object] = {}
    xval: Dict[int, Fraction] = {}         # exact value of floating (and integer) constants
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

    def make_const(value, realm, ct, om_id, ex: Optional[Fraction] = None) -> int:
        v = _np_scalar(value, ct) if ct is not CType.Bool else bool(value)
        if exact and ct is not CType.Bool and ex is None:
            ex = Fraction(int(v)) if ct is CType.Int else Fraction(float(v))
        if exact and ct.is_float:
            # constants that round alike but differ exactly stay apart
            nid = emit("imm", (), realm, ct, (v, ex), om_id)
            nodes[nid].payload = v
        else:
            nid = emit("imm", (), realm, ct, v, om_id)
        const[nid] = v
        if ex is not None:
            xval[nid] = ex
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
            ex = None
            if exact and ct.is_float and getattr(inst, "exact", None) is not None:
                ex = Fraction(inst.exact)
            nid = make_const(inst.value, realm, ct, v, ex)
        elif isinstance(inst, LoadIndex):
            nid = emit("index", (), realm, ct, inst.axis, v)
            if cyclic[inst.axis]:
                rng[nid] = _Rng(inst.axis, (0, 0), (1, -1))
        elif isinstance(inst, LoadSize):
            nid = emit("size", (), realm, ct, inst.axis, v)
            rng[nid] = _Rng(inst.axis, (1, 0), (1, 0))
        elif isinstance(inst, Shift):
            if args[0] in const:
                nid = make_const(const[args[0]], realm, ct, v, xval.get(args[0]))
            elif all(x == 0 for x in inst.vec):
                nid = args[0]
                nodes[nid].om_nodes.append(v)
            else:
                nid = emit("shift", args, realm, ct, inst.vec, v)
        elif isinstance(inst, Broadcast):
            if args[0] in const:
                nid = make_const(const[args[0]], realm, ct, v, xval.get(args[0]))
            else:
                nid = emit("bcast", args, realm, ct, None, v)
                if args[0] in rng:
                    rng[nid] = rng[args[0]]
        elif isinstance(inst, Reduce):
            nid = emit("reduce", args, realm, ct, inst.op, v)
        elif isinstance(inst, Arith):
            op = inst.op
            nid = None
            if (exact and ct.is_float and args and all(a in xval for a in args)
                    and all(nodes[a].ctype is not CType.Bool for a in args)):
                ex = _fold_exact(op, [xval[a] for a in args], [nodes[a].ctype for a in args],
                                 inst.cast_to)
                if ex is not None:
                    nid = make_const(round_exact(ex, ct), realm, ct, v, ex)
            if nid is None and all(a in const for a in args):
                r = _fold(op, [const[a] for a in args], [nodes[a].ctype for a in args], ct,
                          inst.cast_to)
                if r is not None:
                    nid = make_const(r, realm, ct, v)
            if nid is None and exact and ct.is_float and op in (Op.Add, Op.Sub, Op.Mul):
                # x * 0, x * 1, x + 0 and x - 0 of finite operands (the sign of a zero
                # result aside): e.g. the pressure term p * delta(i, j) of an off-diagonal
                # momentum flux, Hydro.hs:113-114
                x0, x1 = xval.get(args[0]), xval.get(args[1])
                if op is Op.Mul and 0 in (x0, x1):
                    nid = make_const(0, realm, ct, v, Fraction(0))
                elif op is Op.Mul and (x0 == 1 or x1 == 1):
                    nid = args[1] if x0 == 1 else args[0]
                elif op is Op.Add and (x0 == 0 or x1 == 0):
                    nid = args[1] if x0 == 0 else args[0]
                elif op is Op.Sub and x1 == 0:
                    nid = args[0]
                if nid is not None and nid in (args[0], args[1]):
                    nodes[nid].om_nodes.append(v)
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
    fx = {n: x for n, x in xval.items() if nodes[n].ctype.is_float}
    return _prune(LKernel(om.kernels[ki].name, nodes, stores, om.ndim, fx))


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
    return LKernel(k.name, out, [(s, remap[v]) for s, v in k.stores], k.ndim,
                   {remap[n]: x for n, x in k.exact.items() if n in remap})
