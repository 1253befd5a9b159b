"""Shared helpers for the parity tests: oracle machines for the current programs."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from paraiso_amd import annotation as A                     # noqa: E402
from paraiso_amd.generator.native import Language, Setup    # noqa: E402
from paraiso_amd.generator.plan import translate            # noqa: E402

CYC2 = (A.Condition.Cyclic,) * 2
CYC3 = (A.Condition.Cyclic,) * 3


def oracle_machine(om_factory, local_size, boundary=None, openmp=True, opt="-O2"):
    """The CPU oracle (PlanTrans restatement) of a program at a compile-time size."""
    from oracle.runner import emitted_machine
    bnd = boundary or (A.Condition.Cyclic,) * len(local_size)
    plan = translate(Setup(tuple(local_size), Language.CPlusPlus, boundary=bnd), om_factory())
    return emitted_machine(plan, openmp=openmp, opt=opt)


def rel_l1(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))
