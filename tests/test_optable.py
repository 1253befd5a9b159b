"""The oracle's operator table (its restatement of PlanTrans.hs:670-711, built with g++)
against each operator's closed form in numpy: exact for IEEE arithmetic, comparisons,
selects, casts and C integer division; within 1 ulp for libm functions (the same double
functions on both sides, FP32 operands widened as the reference's unqualified calls
widen them).  This pins the oracle the GPU operator tests compare against."""
import numpy as np
import pytest

from helpers import oracle_machine
from optable_check import TRANSCENDENTAL, closed_form, ulp_distance
from paraiso_amd.om.dyn import CType
from paraiso_amd.schemes import optable as T

SHAPE = (48, 64)


@pytest.mark.parametrize("real", [CType.Float, CType.Double, CType.Int])
def test_oracle_operator_table(real):
    m = oracle_machine(lambda: T.optable_om(real, signum=False), SHAPE[::-1])
    ins = T.inputs(real, SHAPE, 5)
    for k, v in ins.items():
        m.set(k, v)
    m.run("apply")
    ops = T.INT_OPS if real is CType.Int else T.FLOAT_OPS
    for op in ops:
        got, want = m.get("out_" + op), closed_form(op, real, ins)
        if op in TRANSCENDENTAL:
            assert ulp_distance(got, want).max() <= 1, op
        else:
            assert np.array_equal(got, want), op
    m.close()


def test_pow_is_the_rational_power_operator():
    """``x ^/ y`` builds Arith Pow with the exponent as an immediate (Internal.hs:391)."""
    from paraiso_amd.om import builder as B
    from paraiso_amd.om.arith import Op
    from paraiso_amd.om.dyn import Realm
    from paraiso_amd.om.graph import Arith, Imm
    st = B.static("f", Realm.Array, CType.Double)
    om = B.make_om("P", [], [st], [("k", lambda: B.store(st, B.load(st).pow(T.POW_EXPONENT)))],
                   ndim=1)
    g = om.kernels[0].graph
    ops = [n.inst for n in g.nodes if isinstance(n.inst, Arith)]
    imms = [n.inst for n in g.nodes if isinstance(n.inst, Imm)]
    assert [a.op for a in ops] == [Op.Pow]
    assert [float(i.value) for i in imms] == [1.7]
