"""Conway's Game of Life as an OM program.

``life_om`` restates ``examples/Life/Generator.hs:39-115`` (statics cell,
population, generation; kernels ``init`` and ``proceed``).  ``life_om_legacy``
restates the older-API program ``examples-old/Life-exampled/LifeMain.hs:33-123``
whose generated code and OM dump ship with the reference.
"""
from __future__ import annotations

from ..om import builder as B
from ..om.arith import ReduceOp
from ..om.dyn import CType, Realm

# examples/Life/Generator.hs:79-82
ADJ_VECS = list(zip([-1, 0, 1, -1, 1, -1, 0, 1],
                    [-1, -1, -1, 0, 0, 1, 1, 1]))

# examples-old/Life-exampled/LifeMain.hs:48-51 (R-pentomino)
R5MINO = list(zip([1, 2, 0, 1, 1], [0, 0, 1, 1, 2]))


def _rule(old_cell, num):
    # (c == 0 && n == 3) || (c == 1 && (n >= 2 && n <= 3)); && and || are infixr
    return (old_cell.eq(0) & num.eq(3)) | (old_cell.eq(1) & (num.ge(2) & num.le(3)))


def life_om(name: str = "Life"):
    cell = B.static("cell", Realm.Array, CType.Int)
    population = B.static("population", Realm.Scalar, CType.Int)
    generation = B.static("generation", Realm.Scalar, CType.Int)

    def init():
        B.store(cell, 0)
        B.store(population, 0)
        B.store(generation, 0)

    def proceed():
        old_cell = B.bind(B.load(cell))
        gen = B.bind(B.load(generation))
        neighbours = [B.bind(B.shift(v, old_cell)) for v in ADJ_VECS]
        num = B.bind(B.sum_(neighbours))
        is_alive = B.bind(_rule(old_cell, num))
        new_cell = B.bind(B.select(is_alive, 1, 0))
        B.store(population, B.reduce(ReduceOp.Sum, new_cell))
        B.store(generation, gen + 1)
        B.store(cell, new_cell)

    return B.make_om(name, [], [cell, population, generation],
                     [("init", init), ("proceed", proceed)], ndim=2)


def life_om_legacy(name: str = "Life"):
    population = B.static("population", Realm.Scalar, CType.Int)
    generation = B.static("generation", Realm.Scalar, CType.Int)
    cell = B.static("cell", Realm.Array, CType.Int)

    def proceed():
        c = B.bind(B.load(cell))
        gen = B.bind(B.load(generation))
        shifted = [B.shift(v, c).run() for v in ADJ_VECS]     # forM, then bind each
        neighbours = [B._ret(v) for v in shifted]
        num = B.bind(B.foldl1(lambda a, b: a + b, neighbours))
        is_alive = B.bind(_rule(c, num))
        new_cell = B.bind(B.select(is_alive, B.imm(1, CType.Int, Realm.Array), 0))
        B.store(population, B.reduce(ReduceOp.Sum, new_cell))
        B.store(generation, gen + 1)
        B.store(cell, new_cell)

    def init():
        coord = [B.bind(B.load_index(ax, CType.Int)) for ax in range(2)]
        size = [B.bind(B.load_size(ax, CType.Int, Realm.Array)) for ax in range(2)]
        half = [B.bind(size[ax] // 2) for ax in range(2)]
        rel = [coord[ax] - half[ax] for ax in range(2)]   # Vec (-): unbound

        def agree(point):
            return B.foldl1(lambda a, b: a & b,
                            [rel[i].eq(B.imm(point[i])) for i in range(2)])
        alive = B.bind(B.foldl1(lambda a, b: a | b, [agree(p) for p in R5MINO]))
        c = B.bind(B.select(alive, B.imm(1, CType.Int, Realm.Array), 0))
        B.store(cell, c)
        B.store(population, B.reduce(ReduceOp.Sum, c))
        B.store(generation, B.imm(0, CType.Int, Realm.Scalar))

    return B.make_om(name, [], [population, generation, cell],
                     [("init", init), ("proceed", proceed)], ndim=2)
