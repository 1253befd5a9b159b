"""OM-level transforms used by the HIP backend's tuner (no counterpart in the reference).

``unroll(om, kernel, k)`` adds a kernel ``<kernel>_x<k>`` equal to ``k``
consecutive calls of ``kernel``: copy c+1 reads, instead of each static, the
value copy c stored to it (the reference's store semantics: every load of a
call sees the statics as they were before the call), and the last copy's
values are stored.  Values of earlier copies that only fed superseded stores
die in dead-code elimination.  Lowered by the tile compiler, one launch then
advances several steps with a correspondingly deeper halo: temporal blocking,
traded against recomputation in the halo.  A Reduce whose result feeds a
later copy's arrays (a CFL step) still needs a global synchronisation, so
such kernels gain nothing and keep one launch per level.
"""
from __future__ import annotations

from .adjoint import _keep_anot, _rebuild, _replay_value
from .om import builder as B
from .om.graph import OM, Load, Store


def unroll(om: OM, kernel: str, k: int, name: str = None) -> OM:
    if k < 1:
        raise ValueError("unroll factor must be positive")
    fwd = om.kernels[om.kernel_index(kernel)].graph
    statics = list(om.setup.static_values)

    def body():
        current = {}                 # static index -> bound builder of its latest value
        for _copy in range(k):
            g = fwd
            mapped = {}
            stored = {}
            for v, node in enumerate(g.nodes):
                if node.is_value:
                    inst_id, inst = g.inst_of_value(v)
                    if isinstance(inst, Load) and inst.static in current:
                        mapped[v] = current[inst.static]
                    else:
                        mapped[v] = B.bind(_keep_anot(node, _replay_value(g, v, mapped, statics)))
                elif isinstance(node.inst, Store):
                    stored[node.inst.static] = mapped[g.pre(v)[0]]
            current.update(stored)
        for s, val in current.items():
            B.store(statics[s], val)

    kernels = [(kk.name, _rebuild(kk, statics)) for kk in om.kernels]
    kernels.append((name or "%s_x%d" % (kernel, k), body))
    return B.make_om(om.name, list(om.setup.global_anot), statics, kernels, om.ndim)
