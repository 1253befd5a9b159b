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
from .om.dyn import Realm
from .om.graph import OM, Load, Reduce, Store


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


def pipeline_reduce(om: OM, kernel: str, name: str = None, prime: str = None) -> OM:
    """Carry a kernel's reductions from one call to the next.

    For a kernel whose Reduce instructions read only the statics as the call
    finds them (level-0 reductions, such as Hydro's CFL minimum), this adds

    * ``<kernel>_prime``: evaluates the reductions on the current statics and
      stores each result into a carried scalar static ``__carry_<kernel>_<i>``;
    * ``<kernel>_pipelined``: the kernel with each Reduce result replaced by its
      carried value, followed by the reductions evaluated on the values this call
      stores, which go back into the carried statics.

    ``prime`` once and then ``pipelined`` n times stores exactly what ``kernel``
    n times stores (the same operations on the same values), but every call is a
    single pass over the grid: the reductions of the next call ride on the
    store phase of this one instead of re-reading the state.
    """
    ki = om.kernel_index(kernel)
    g = om.kernels[ki].graph
    statics = list(om.setup.static_values)
    reduces = [v for v, nd in enumerate(g.nodes)
               if nd.is_value and isinstance(g.inst_of_value(v)[1], Reduce)]
    if not reduces:
        raise ValueError("kernel %s has no Reduce to carry" % kernel)
    # the cone of the reductions: values they depend on
    cone = set()
    stack = list(reduces)
    while stack:
        v = stack.pop()
        if v in cone:
            continue
        cone.add(v)
        inst_id, inst = g.inst_of_value(v)
        for a in g.pre(inst_id):
            stack.append(a)
    for r in reduces:
        inst_id, _ = g.inst_of_value(r)
        upstream = set()
        st = list(g.pre(inst_id))
        while st:
            v = st.pop()
            if v in upstream:
                continue
            upstream.add(v)
            iid, inst = g.inst_of_value(v)
            if isinstance(inst, Reduce):
                raise ValueError("a Reduce of %s depends on another Reduce" % kernel)
            st.extend(g.pre(iid))
    carry = {r: B.static("__carry_%s_%d" % (kernel, i), Realm.Scalar, g.nodes[r].dyn.ctype)
             for i, r in enumerate(reduces)}
    all_statics = statics + [carry[r] for r in reduces]

    def replay_cone(current):
        mapped = {}
        for v in sorted(cone):
            node = g.nodes[v]
            inst_id, inst = g.inst_of_value(v)
            if isinstance(inst, Load) and inst.static in current:
                mapped[v] = current[inst.static]
            else:
                mapped[v] = B.bind(_keep_anot(node, _replay_value(g, v, mapped, all_statics)))
        return mapped

    def prime_body():
        mapped = replay_cone({})
        for r in reduces:
            B.store(carry[r], mapped[r])

    def pipelined_body():
        mapped, stored = {}, {}
        for v, node in enumerate(g.nodes):
            if node.is_value:
                inst_id, inst = g.inst_of_value(v)
                if isinstance(inst, Reduce):
                    mapped[v] = B.bind(B.load(carry[v]))
                else:
                    mapped[v] = B.bind(_keep_anot(node, _replay_value(g, v, mapped, all_statics)))
            elif isinstance(node.inst, Store):
                stored[node.inst.static] = mapped[g.pre(v)[0]]
        nxt = replay_cone(stored)
        for s, val in stored.items():
            B.store(statics[s], val)
        for r in reduces:
            B.store(carry[r], nxt[r])

    kernels = [(kk.name, _rebuild(kk, all_statics)) for kk in om.kernels]
    kernels.append((prime or "%s_prime" % kernel, prime_body))
    kernels.append((name or "%s_pipelined" % kernel, pipelined_body))
    return B.make_om(om.name, list(om.setup.global_anot), all_statics, kernels, om.ndim)
