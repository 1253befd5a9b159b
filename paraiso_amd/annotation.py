"""Typed annotation bags attached to every OM node.

Semantics follow ``Language/Paraiso/Annotation.hs:17-60``: an annotation is a
list of typed items; ``set`` replaces every item of the same type, ``weak_set``
only adds when no item of that type exists, ``to_maybe`` returns the first.
The item types follow ``Annotation/Allocation.hs``, ``Annotation/Boundary.hs``,
``Annotation/Dependency.hs``, ``Annotation/Execution.hs`` and
``Annotation/SyncThreads.hs``.
"""
from __future__ import annotations

from dataclasses import dataclass
from enum import Enum
from typing import Optional, Tuple


def add(x, anot: list) -> list:
    return [x] + list(anot)


def set_(x, anot: list) -> list:
    return [x] + [a for a in anot if type(a) is not type(x)]


def weak_set(x, anot: list) -> list:
    if any(type(a) is type(x) for a in anot):
        return list(anot)
    return [x] + list(anot)


def to_list(cls, anot: list) -> list:
    return [a for a in anot if type(a) is cls]


def to_maybe(cls, anot: list):
    for a in anot:
        if type(a) is cls:
            return a
    return None


# ---- Allocation (Annotation/Allocation.hs:14-21)
class Allocation(Enum):
    Existing = "Existing"
    Manifest = "Manifest"
    Delayed = "Delayed"


@dataclass(frozen=True)
class AllocationChoice:
    choices: Tuple[Allocation, ...]


# ---- Boundary (Annotation/Boundary.hs:21-38)
class Condition(Enum):
    Open = "Open"
    Cyclic = "Cyclic"


class NearBoundary:
    """``NegaInfinity | LowerBoundary a | UpperBoundary a | PosiInfinity``.

    The derived ``Ord`` compares the constructor first, then the payload.
    """
    __slots__ = ("tag", "x")
    NEG, LOW, UP, POS = 0, 1, 2, 3

    def __init__(self, tag: int, x: int = 0):
        self.tag = tag
        self.x = x if tag in (1, 2) else 0

    def key(self):
        return (self.tag, self.x)

    def __eq__(self, o):
        return isinstance(o, NearBoundary) and self.key() == o.key()

    def __hash__(self):
        return hash(self.key())

    def __lt__(self, o):
        return self.key() < o.key()

    def __le__(self, o):
        return self.key() <= o.key()

    def shifted(self, d: int) -> "NearBoundary":
        if self.tag in (1, 2):
            return NearBoundary(self.tag, self.x + d)
        return self

    def pp(self, lower: bool) -> str:
        # OM/PrettyPrint.hs:101-104
        if self.tag == 0:
            return "[-inf"
        if self.tag == 3:
            return "+inf]"
        if self.tag == 1:
            return "[" + str(self.x)
        return str(self.x) + "]"

    def __repr__(self):
        return ["NegaInfinity", "LowerBoundary %d" % self.x,
                "UpperBoundary %d" % self.x, "PosiInfinity"][self.tag]


NEG_INF = NearBoundary(NearBoundary.NEG)
POS_INF = NearBoundary(NearBoundary.POS)


class Interval:
    """Half-open interval of NearBoundary values (``Interval.hs:18-33``)."""
    __slots__ = ("lower", "upper")

    def __init__(self, lower: Optional[NearBoundary], upper: Optional[NearBoundary]):
        self.lower = lower
        self.upper = upper

    @property
    def empty(self) -> bool:
        return self.lower is None

    def intersection(self, o: "Interval") -> "Interval":
        if self.empty or o.empty:
            return EMPTY
        lo = max(self.lower, o.lower)
        up = min(self.upper, o.upper)
        if lo >= up:
            return EMPTY
        return Interval(lo, up)

    def __eq__(self, o):
        return isinstance(o, Interval) and self.lower == o.lower and self.upper == o.upper

    def __hash__(self):
        return hash((self.lower, self.upper))

    def pp(self) -> str:
        if self.empty:
            return "[empty]"
        return self.lower.pp(True) + ".." + self.upper.pp(False)

    def __repr__(self):
        return "Interval(%r, %r)" % (self.lower, self.upper)


EMPTY = Interval(None, None)


@dataclass(frozen=True)
class Valid:
    intervals: Tuple[Interval, ...]

    def intersection(self, o: "Valid") -> "Valid":
        if len(self.intervals) != len(o.intervals):
            raise ValueError("length mismatch in merging two Valid")
        return Valid(tuple(a.intersection(b) for a, b in zip(self.intervals, o.intervals)))

    def pp(self) -> str:
        return " ".join(i.pp() for i in self.intervals)


def full_valid(ndim: int) -> Valid:
    return Valid(tuple(Interval(NearBoundary(1, 0), NearBoundary(2, 0)) for _ in range(ndim)))


def infinite_valid(ndim: int) -> Valid:
    return Valid(tuple(Interval(NEG_INF, POS_INF) for _ in range(ndim)))


# ---- Dependency (Annotation/Dependency.hs:23-51)
@dataclass(frozen=True)
class Direct:
    nodes: Tuple[int, ...]


@dataclass(frozen=True)
class Indirect:
    nodes: Tuple[int, ...]


@dataclass(frozen=True)
class Calc:
    nodes: frozenset


@dataclass(frozen=True)
class KernelWriteGroup:
    gid: int


@dataclass(frozen=True)
class OMWriteGroup:
    gid: int


# ---- Execution (Annotation/Execution.hs:14-15)
@dataclass(frozen=True)
class Alive:
    alive: bool


# ---- SyncThreads (Annotation/SyncThreads.hs:13-14)
class Timing(Enum):
    Pre = "Pre"
    Post = "Post"


# ---- optimisation level recorded in the global annotation (Optimization.hs:55-61)
class Level(Enum):
    Unoptimized = 0
    O0 = 1
    O1 = 2
    O2 = 3
    O3 = 4
