"""The Orthotope Machine dataflow graph.

Follows ``Language/Paraiso/OM/Graph.hs:40-135`` and ``Language/Paraiso/OM.hs:16-41``:
a bipartite DAG of value nodes (``NValue DynValue anot``) and instruction
nodes (``NInst Inst anot``); edges carry the argument position (``EOrd i``).
Node ids are dense and equal creation order (``Builder/Internal.hs:106-119``).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Tuple

from .arith import ARITY, Op, ReduceOp
from .dyn import CType, DynValue


# ---------------------------------------------------------------- instructions
@dataclass(frozen=True)
class Load:
    static: int

    def show(self):
        return "Load static[%d]" % self.static


@dataclass(frozen=True)
class Store:
    static: int

    def show(self):
        return "Store static[%d]" % self.static


@dataclass(frozen=True)
class Reduce:
    op: ReduceOp

    def show(self):
        return "Reduce %s" % self.op.value


@dataclass(frozen=True)
class Broadcast:
    def show(self):
        return "Broadcast"


@dataclass(frozen=True)
class LoadIndex:
    axis: int

    def show(self):
        return "LoadIndex (Axis {axisIndex = %d})" % self.axis


@dataclass(frozen=True)
class LoadSize:
    axis: int

    def show(self):
        return "LoadSize (Axis {axisIndex = %d})" % self.axis


def show_vec(v: Tuple[int, ...]) -> str:
    s = "Vec"
    for x in v:
        s = "(%s :~ %d)" % (s, x) if s != "Vec" else "(Vec :~ %d)" % x
    return s


@dataclass(frozen=True)
class Shift:
    vec: Tuple[int, ...]

    def show(self):
        return "Shift %s" % show_vec(self.vec)


@dataclass(frozen=True)
class Imm:
    """An immediate; ``value`` is already converted to ``ctype`` (resolved late).

    ``exact`` keeps the literal the program wrote (an int or a Fraction) before that
    rounding; it takes no part in equality or in the printed form."""
    value: object
    ctype: Optional[CType] = None
    exact: object = field(default=None, compare=False)

    def show(self, old: bool = False):
        if old:
            return "Imm <<%s>>" % (self.ctype.value if self.ctype else "?")
        from .literal import c_literal
        return "Imm " + c_literal(self.value, self.ctype)


@dataclass(frozen=True)
class Arith:
    op: Op
    cast_to: Optional[CType] = None   # only for Op.Cast

    def show(self):
        if self.op is Op.Cast:
            return "Arith (Cast %s)" % self.cast_to.value
        return "Arith %s" % self.op.value


def inst_arity(inst) -> Tuple[int, int]:
    if isinstance(inst, Load):
        return (0, 1)
    if isinstance(inst, Store):
        return (1, 0)
    if isinstance(inst, (Reduce, Broadcast, Shift)):
        return (1, 1)
    if isinstance(inst, (LoadIndex, LoadSize, Imm)):
        return (0, 1)
    if isinstance(inst, Arith):
        return ARITY[inst.op]
    raise TypeError(inst)


# ---------------------------------------------------------------- nodes
class Node:
    """A value node (``dyn`` set) or an instruction node (``inst`` set)."""
    __slots__ = ("dyn", "inst", "anot", "ty")

    def __init__(self, dyn=None, inst=None, anot=None, ty=None):
        self.dyn: Optional[DynValue] = dyn
        self.inst = inst
        self.anot: list = list(anot or [])
        self.ty = ty          # (realm TyVar, content TyVar) while building

    @property
    def is_value(self) -> bool:
        return self.inst is None

    def copy(self) -> "Node":
        return Node(self.dyn, self.inst, list(self.anot), None)


class Graph:
    """Dense DAG: ``preds[i]`` is the argument-ordered predecessor list."""

    def __init__(self):
        self.nodes: List[Node] = []
        self.preds: List[List[int]] = []
        self._succs: Optional[List[List[Tuple[int, int]]]] = None

    def add(self, preds: List[int], node: Node) -> int:
        n = len(self.nodes)
        self.nodes.append(node)
        self.preds.append(list(preds))
        self._succs = None
        return n

    def __len__(self):
        return len(self.nodes)

    def succ_edges(self, i: int) -> List[Tuple[int, int]]:
        """(successor id, argument position) pairs."""
        if self._succs is None:
            s = [[] for _ in self.nodes]
            for j, ps in enumerate(self.preds):
                for pos, p in enumerate(ps):
                    s[p].append((j, pos))
            self._succs = s
        return self._succs[i]

    def succ(self, i: int) -> List[int]:
        return [j for j, _ in self.succ_edges(i)]

    def pre(self, i: int) -> List[int]:
        return self.preds[i]

    def copy(self) -> "Graph":
        g = Graph()
        g.nodes = [n.copy() for n in self.nodes]
        g.preds = [list(p) for p in self.preds]
        return g

    def inst_of_value(self, v: int):
        """The instruction node producing value node ``v`` (its single pred)."""
        ps = self.preds[v]
        if len(ps) != 1:
            raise ValueError("value node %d has %d preds" % (v, len(ps)))
        return ps[0], self.nodes[ps[0]].inst


@dataclass
class Kernel:
    name: str
    graph: Graph


@dataclass
class Named:
    name: str
    dyn: DynValue


@dataclass
class Setup:
    """OM setup: the static variables and the global annotation (Graph.hs:62-70)."""
    static_values: List[Named]
    global_anot: list = field(default_factory=list)


@dataclass
class OM:
    name: str
    setup: Setup
    kernels: List[Kernel]
    ndim: int

    def static_index(self, name: str) -> int:
        hits = [i for i, s in enumerate(self.setup.static_values) if s.name == name]
        if len(hits) != 1:
            raise KeyError("%d match found for '%s'" % (len(hits), name))
        return hits[0]

    def kernel_index(self, name: str) -> int:
        for i, k in enumerate(self.kernels):
            if k.name == name:
                return i
        raise KeyError(name)
