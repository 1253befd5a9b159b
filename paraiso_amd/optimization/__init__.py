"""Mandatory analyses run on every OM before code generation.

Pipeline (``Language/Paraiso/Optimization.hs:28-53``):
dead-code elimination -> allocation decision -> boundary (Valid) analysis ->
dependency analysis + write grouping -> identity.  Only level O0 does work;
higher levels recurse to O0, and a level already recorded in the global
annotation is not redone (``Optimization.hs:33-45``).
"""
from __future__ import annotations

import copy

from .. import annotation as A
from ..om.graph import OM, Kernel
from .allocation import decide_allocation
from .boundary import boundary_analysis
from .dce import dead_code_elimination
from .dependency import write_grouping


def _gmap(f, om: OM) -> OM:
    return OM(om.name, om.setup, [Kernel(k.name, f(k.graph, om)) for k in om.kernels], om.ndim)


def optimize(level: A.Level, om: OM) -> OM:
    old = A.to_maybe(A.Level, om.setup.global_anot)
    if old is not None and level.value <= old.value:
        return om
    om1 = _gmap(lambda g, o: dead_code_elimination(g), om)
    om1 = _gmap(lambda g, o: decide_allocation(g), om1)
    om1 = _gmap(lambda g, o: boundary_analysis(g, o.ndim), om1)
    om1 = write_grouping(om1)
    setup = copy.copy(om1.setup)
    setup.global_anot = A.set_(level, setup.global_anot)
    return OM(om1.name, setup, om1.kernels, om1.ndim)
