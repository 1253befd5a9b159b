"""ORACLE -- test infrastructure only: compile the CPU machines the GPU parity
tests use, so the GPU box does not spend its time in g++."""
import os
import sys
from concurrent.futures import ThreadPoolExecutor
from fractions import Fraction

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from paraiso_amd import annotation as A                      # noqa: E402
from paraiso_amd.generator.native import Language, Setup     # noqa: E402
from paraiso_amd.generator.plan import translate             # noqa: E402
from paraiso_amd.om.dyn import CType                         # noqa: E402

from . import cpp_emitter                                    # noqa: E402

G = Fraction(7, 5)


def _cases():
    from paraiso_amd.schemes.life import life_om
    from paraiso_amd.schemes.hydro import hydro_om
    cyc2 = (A.Condition.Cyclic,) * 2
    out = []
    for n in [(64, 64), (100, 37), (8, 8), (256, 128), (1024, 1024), (96, 96), (256, 192)]:
        out.append((life_om, n, cyc2))
    out.append((lambda: hydro_om(CType.Float, G), (2048, 2048), cyc2))   # bench CPU baseline
    open2 = (A.Condition.Open,) * 2
    for n in [(64, 64), (100, 37), (9, 130)]:
        out.append((life_om, n, open2))
    from paraiso_amd.adjoint import add_adjoint
    from paraiso_amd.schemes.linear_wave import linear_wave_om
    cyc3 = (A.Condition.Cyclic,) * 3
    out.append((lambda: add_adjoint(linear_wave_om(3, CType.Double), "proceed",
                                    ["fieldF", "fieldG"]), (16, 12, 10), cyc3))
    out.append((lambda: add_adjoint(linear_wave_om(3, CType.Double), "proceed",
                                    ["fieldF", "fieldG"]), (40, 9, 70), cyc3))
    out.append((lambda: add_adjoint(linear_wave_om(3, CType.Double), "proceed",
                                    ["fieldF", "fieldG"]), (64, 64, 64), cyc3))
    out.append((lambda: add_adjoint(linear_wave_om(3, CType.Float), "proceed",
                                    ["fieldF", "fieldG"]), (64, 64, 64), cyc3))
    for n in [(64, 48), (96, 64), (48, 32), (80, 56), (512, 512)]:
        out.append((lambda: hydro_om(CType.Double, G), n, cyc2))
        out.append((lambda: hydro_om(CType.Float, G), n, cyc2))
    out.append((lambda: hydro_om(CType.Float, G), (1088, 544), cyc2))
    out.append((lambda: hydro_om(CType.Double, G), (544, 544), cyc2))
    return out


def prebuild(jobs: int = 4):
    plans = [translate(Setup(n, Language.CPlusPlus, boundary=b), f()) for f, n, b in _cases()]
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(lambda p: cpp_emitter.build(p), plans))


if __name__ == "__main__":
    prebuild()
