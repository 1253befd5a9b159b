"""The OM transforms, checked on the CPU oracle (the PlanTrans restatement runs any OM):

* ``pipeline_reduce``: prime + n x pipelined stores what n x proceed stores (Hydro, FP64);
* ``unroll``: proceed_x<k> stores what k x proceed stores (Life x4 and x8, LinearWave x2 and x4);
* ``add_adjoint``: the adjoint kernel of the linear wave step is the transpose of the
  step, <A x, y> = <x, A^T y>, to rounding.
The GPU tests check the same identities on the HIP machines."""
from fractions import Fraction

import numpy as np
import pytest

from helpers import oracle_machine
from paraiso_amd import inputs
from paraiso_amd.adjoint import add_adjoint
from paraiso_amd.om.dyn import CType
from paraiso_amd.schemes.hydro import hydro_om
from paraiso_amd.schemes.life import life_om
from paraiso_amd.schemes.linear_wave import linear_wave_om
from paraiso_amd.transform import pipeline_reduce, unroll

F = ("density", "velocity0", "velocity1", "pressure")


def test_pipelined_hydro_equals_proceed():
    nx, ny = 40, 24
    mk = lambda: pipeline_reduce(hydro_om(CType.Double, Fraction(7, 5)), "proceed")
    a, b = oracle_machine(mk, (nx, ny)), oracle_machine(mk, (nx, ny))
    for m in (a, b):
        for k, v in inputs.hydro_scalars(nx, ny).items():
            m.set(k, v)
        for name, arr in zip(F, inputs.kh_init(nx, ny, 6)):
            m.set(name, arr)
    a.run("proceed", 7)
    b.run("proceed_prime")
    b.run("proceed_pipelined", 7)
    assert a.get("time") == b.get("time")
    for f in F:
        assert np.array_equal(a.get(f), b.get(f)), f


@pytest.mark.parametrize("k", [4, 8])
def test_unrolled_life_equals_k_generations(k):
    nx, ny = 30, 22
    mk = lambda: unroll(life_om(), "proceed", k)
    a, b = oracle_machine(mk, (nx, ny)), oracle_machine(mk, (nx, ny))
    cells = inputs.life_init((ny, nx), 5)
    for m in (a, b):
        m.run("init")
        m.set("cell", cells)
    a.run("proceed", 3 * k)
    b.run("proceed_x%d" % k, 3)
    assert np.array_equal(a.get("cell"), b.get("cell"))
    assert a.get("population") == b.get("population")
    assert a.get("generation") == b.get("generation") == 3 * k


def wave(n, real=CType.Double, k=2):
    m = oracle_machine(lambda: unroll(add_adjoint(linear_wave_om(3, real), "proceed",
                                                  ["fieldF", "fieldG"]), "proceed", k), n)
    m.set("c", 1.0)
    m.set("dx", 1.0)
    m.set("dt", 0.5)
    return m


@pytest.mark.parametrize("k", [2, 4])
def test_unrolled_wave_equals_k_steps(k):
    n = (12, 10, 8)
    rng = np.random.default_rng(2)
    f0, g0 = rng.standard_normal(n[::-1]), rng.standard_normal(n[::-1])
    a, b = wave(n, k=k), wave(n, k=k)
    for m in (a, b):
        m.set("fieldF", f0)
        m.set("fieldG", g0)
    a.run("proceed", 3 * k)
    b.run("proceed_x%d" % k, 3)
    for f in ("fieldF", "fieldG"):
        assert np.array_equal(a.get(f), b.get(f)), f
    assert a.get("energy") == b.get("energy")


def test_wave_adjoint_is_the_transpose():
    n = (12, 10, 8)
    rng = np.random.default_rng(4)
    x = [rng.standard_normal(n[::-1]) for _ in range(2)]
    y = [rng.standard_normal(n[::-1]) for _ in range(2)]
    m = wave(n)
    m.set("fieldF", x[0])
    m.set("fieldG", x[1])
    m.run("proceed")
    ax = [m.get("fieldF"), m.get("fieldG")]
    m.set("adj_fieldF", y[0])
    m.set("adj_fieldG", y[1])
    m.run("proceed_adjoint")
    aty = [m.get("adj_fieldF"), m.get("adj_fieldG")]
    lhs = sum(float((p * q).sum()) for p, q in zip(ax, y))
    rhs = sum(float((p * q).sum()) for p, q in zip(x, aty))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
