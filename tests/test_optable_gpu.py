"""Every operator the HIP emitter lowers (paraiso_amd/schemes/optable.py), one kernel per
type, on the GPU against the oracle on the same seeded inputs:

* correctly rounded builds: bit-exact for IEEE arithmetic (+ - * /, 1/x, sqrt), min, max,
  abs, neg, selects, comparisons, logic, casts and C integer division and remainder;
* libm functions (exp, log, sin, cos, tan, asin, acos, atan, atan2, pow): the device's
  double functions against glibc's, within 1 ulp for FP32 operands (computed in double,
  as the reference's unqualified calls do, then rounded) and 2 ulp for FP64;
* the fast FP32 build (hardware reciprocal and square root): within 2 ulp for division,
  reciprocal and square root, the rest as above.
Signum, which the reference prints as a call to an undefined function, is checked against
its definition (numpy.sign)."""
import numpy as np
import pytest

from helpers import oracle_machine
from optable_check import TRANSCENDENTAL, ulp_distance
from paraiso_amd import variants
from paraiso_amd.om.dyn import CType
from paraiso_amd.schemes import optable as T

pytestmark = pytest.mark.gpu
SHAPE = (48, 64)
ULP_LIBM = {CType.Float: 1, CType.Double: 2}


@pytest.mark.parametrize("variant,real", [("optable_f64", CType.Double),
                                          ("optable_f32", CType.Float),
                                          ("optable_f32_fast", CType.Float),
                                          ("optable_i32", CType.Int)])
def test_gpu_operator_table(require_gpu, variant, real):
    print("operator table %s" % variant, flush=True)
    gpu = variants.machine(variant, SHAPE[::-1])
    ora = oracle_machine(lambda: T.optable_om(real, signum=False), SHAPE[::-1])
    ins = T.inputs(real, SHAPE, 5)
    for m in (gpu, ora):
        for k, v in ins.items():
            m.set(k, v)
        m.run("apply")
    fast = variant.endswith("_fast")
    seen = {}
    for op in (T.INT_OPS if real is CType.Int else T.FLOAT_OPS):
        g, o = gpu.get("out_" + op), ora.get("out_" + op)
        if op in TRANSCENDENTAL:
            seen[op] = ulp_distance(g, o).max()
            assert seen[op] <= ULP_LIBM[real], (op, seen[op])
        elif fast and op in ("div", "inv", "sqrt"):
            seen[op] = ulp_distance(g, o).max()
            assert seen[op] <= 2, (op, seen[op])
        else:
            assert np.array_equal(g, o), op
    print("largest ulp distances: %s" % {k: int(v) for k, v in seen.items()})


@pytest.mark.parametrize("variant,real", [("optable_f64", CType.Double),
                                          ("optable_f32", CType.Float)])
def test_gpu_signum(require_gpu, variant, real):
    gpu = variants.machine(variant, SHAPE[::-1])
    ins = T.inputs(real, SHAPE, 6)
    gpu.set("x", ins["x"])
    gpu.run("signum")
    got = gpu.get("out_signum")
    assert (ins["x"] == 0).any() and np.array_equal(got, np.sign(ins["x"]))
