"""Reverse-mode transform of an OM kernel (no counterpart in the reference).

``add_adjoint(om, kernel, wrt)`` returns a machine with one more kernel,
``<kernel>_adjoint``, and one more Array static ``adj_<s>`` per differentiated
static ``s``.  Running the adjoint kernel with the primal statics holding the
state u_n *before* a forward step and ``adj_*`` holding dJ/du_{n+1} leaves
dJ/du_n in ``adj_*``: one vector-Jacobian product of the step.

The adjoint kernel is an ordinary OM kernel built with the Builder:
1. the forward kernel is replayed node for node (so the primal values of the
   step are recomputed from u_n inside the same fused launch),
2. every Store of a differentiated static seeds the stored value's cotangent
   with ``load adj_<s>``,
3. value nodes are visited in reverse creation order; each complete
   cotangent is bound once and pushed to the arguments:
   Add/Sub/Neg/Mul/Div/Inv/Sqrt by their derivatives, Shift v by Shift -v,
   Select/Max/Min route the cotangent to the branch the forward chose
   (``std::max``/``std::min`` tie rules), Abs by the sign, Cast by the cast back,
4. the cotangents reaching the Loads of each differentiated static are summed
   and stored into ``adj_<s>``; for a static the kernel reads but does not store,
   the cotangent it had after the kernel is added (``adj_<s>`` is loaded), unless
   the caller names it in ``zero_in`` (known to be zero there).
Scalars (dt, dR, ...) and Reduce results are treated as constants: the
transform differentiates the array state only, which is what a fixed-dt step
needs.  A Reduce feeding the differentiated state (a CFL dt) is non-smooth
and is rejected.
"""
from __future__ import annotations

from typing import Dict, List, Sequence

from .om import builder as B
from .om.arith import Op
from .om.dyn import Realm
from .om.graph import (OM, Arith, Broadcast, Imm, Load, LoadIndex, LoadSize, Reduce, Shift,
                       Store)


def _replay_value(g, v, mapped, statics):
    inst_id, inst = g.inst_of_value(v)
    args = [mapped[a] for a in g.pre(inst_id)]
    dyn = g.nodes[v].dyn
    if isinstance(inst, Load):
        return B.load(statics[inst.static])
    if isinstance(inst, Imm):
        return B.imm(inst.value if inst.exact is None else inst.exact, inst.ctype, dyn.realm)
    if isinstance(inst, LoadIndex):
        return B.load_index(inst.axis, dyn.ctype)
    if isinstance(inst, LoadSize):
        return B.load_size(inst.axis, dyn.ctype, dyn.realm)
    if isinstance(inst, Shift):
        return B.shift(inst.vec, args[0])
    if isinstance(inst, Broadcast):
        return B.broadcast(args[0])
    if isinstance(inst, Reduce):
        return B.reduce(inst.op, args[0])
    if isinstance(inst, Arith):
        op = inst.op
        if op is Op.Cast:
            return B.cast(args[0], inst.cast_to)
        if op is Op.Select:
            return B.select(*args)
        if op in (Op.EQ, Op.NE, Op.LT, Op.LE, Op.GT, Op.GE):
            return B.mk_op2b(op, *args)
        if len(args) == 1:
            return B.mk_op1(op, args[0])
        if len(args) == 2:
            return B.mk_op2(op, args[0], args[1])
    raise NotImplementedError("cannot replay %r" % (inst,))


def add_adjoint(om: OM, kernel: str, wrt: Sequence[str], name: str = None,
                zero_in: Sequence[str] = ()) -> OM:
    """Add ``<kernel>_adjoint`` (or ``name``): one vector-Jacobian product of ``kernel``
    with respect to the Array statics ``wrt``.  Statics in ``zero_in`` that the kernel
    does not store have a zero cotangent after it (nothing later reads them), so their
    adjoint is not accumulated."""
    fwd = om.kernels[om.kernel_index(kernel)].graph
    statics = list(om.setup.static_values)
    by_name = {s.name: i for i, s in enumerate(statics)}
    adj_statics = []
    for s in wrt:
        st = statics[by_name[s]]
        if st.dyn.realm is not Realm.Array or not st.dyn.ctype.is_float:
            raise ValueError("only floating Array statics can be differentiated: %s" % s)
        adj_statics.append(B.static("adj_" + s, Realm.Array, st.dyn.ctype))
    existing = {s.name for s in statics}
    all_statics = statics + [a for a in adj_statics if a.name not in existing]
    adj_of = {by_name[s]: next(x for x in all_statics if x.name == "adj_" + s) for s in wrt}
    stored = {n.inst.static for n in fwd.nodes if isinstance(n.inst, Store)}
    _check_smooth(fwd, set(adj_of))

    def body():
        g = fwd
        mapped: Dict[int, B.Builder] = {}
        for v, node in enumerate(g.nodes):
            if node.is_value:
                mapped[v] = B.bind(_keep_anot(node, _replay_value(g, v, mapped, all_statics)))
        cot: Dict[int, List[B.Builder]] = {}
        for v, node in enumerate(g.nodes):
            if isinstance(node.inst, Store) and node.inst.static in adj_of:
                val = g.pre(v)[0]
                cot.setdefault(val, []).append(B.load(adj_of[node.inst.static]))
        load_cots: Dict[int, List[B.Builder]] = {}
        values = [v for v, n in enumerate(g.nodes) if n.is_value]
        for v in reversed(values):
            contribs = cot.get(v)
            if not contribs:
                continue
            dyn = g.nodes[v].dyn
            if dyn.realm is not Realm.Array or not dyn.ctype.is_float:
                continue
            ybar = B.bind(B.foldl1(lambda a, b: a + b, contribs))
            inst_id, inst = g.inst_of_value(v)
            args = g.pre(inst_id)
            if isinstance(inst, Load):
                load_cots.setdefault(inst.static, []).append(ybar)
                continue
            _push(g, v, inst, args, ybar, mapped, cot)
        for sidx, adj in adj_of.items():
            cs = list(load_cots.get(sidx, []))
            st = statics[sidx]
            if sidx not in stored and st.name not in zero_in:
                # read, not written: the cotangent after the kernel carries through
                cs.insert(0, B.load(adj))
            if cs:
                B.store(adj, B.foldl1(lambda a, b: a + b, cs))
            else:
                B.store(adj, B.imm(0, st.dyn.ctype, Realm.Array))

    kernels = [(k.name, _rebuild(k, all_statics)) for k in om.kernels]
    kernels.append((name or kernel + "_adjoint", body))
    return B.make_om(om.name, list(om.setup.global_anot), all_statics, kernels, om.ndim)


def _rebuild(k, statics):
    """A builder body that recreates kernel ``k`` node for node."""
    def body():
        g = k.graph
        mapped = {}
        for v, node in enumerate(g.nodes):
            if node.is_value:
                mapped[v] = B.bind(_keep_anot(node, _replay_value(g, v, mapped, statics)))
            elif isinstance(node.inst, Store):
                B.store(statics[node.inst.static], mapped[g.pre(v)[0]])
    return body


def _keep_anot(node, b):
    """Carry the program's own annotations (e.g. Manifest hints) to the replay."""
    src = list(node.anot)
    if not src:
        return b
    return B.annotate(lambda _a: list(src), b)


def _check_smooth(g, adj_store_statics):
    """Reject a Reduce (e.g. a CFL minimum) upstream of a differentiated store."""
    stack = [g.pre(v)[0] for v, n in enumerate(g.nodes)
             if isinstance(n.inst, Store) and n.inst.static in adj_store_statics]
    seen = set()
    while stack:
        x = stack.pop()
        if x in seen:
            continue
        seen.add(x)
        if isinstance(g.nodes[x].inst, Reduce):
            raise ValueError("a Reduce feeds the differentiated state; the step is not "
                             "differentiable (use a fixed dt)")
        stack.extend(g.pre(x))


def _differentiable(g, a) -> bool:
    d = g.nodes[a].dyn
    return d.realm is Realm.Array and d.ctype.is_float


def _push(g, v, inst, args, ybar, mapped, cot):
    def add(a, contrib):
        if _differentiable(g, a):
            cot.setdefault(a, []).append(contrib)

    ct = g.nodes[v].dyn.ctype
    zero = B.imm(0, ct, Realm.Array)
    y = mapped[v]
    if isinstance(inst, Shift):
        add(args[0], B.shift(tuple(-x for x in inst.vec), ybar))
        return
    if isinstance(inst, Broadcast):
        return                      # scalar argument: held constant
    if isinstance(inst, Reduce):
        return
    if not isinstance(inst, Arith):
        return
    op = inst.op
    m = [mapped[a] for a in args]
    if op is Op.Add:
        add(args[0], ybar)
        add(args[1], ybar)
    elif op is Op.Sub:
        add(args[0], ybar)
        add(args[1], -ybar)
    elif op is Op.Neg:
        add(args[0], -ybar)
    elif op is Op.Mul:
        add(args[0], ybar * m[1])
        add(args[1], ybar * m[0])
    elif op is Op.Div:
        add(args[0], ybar / m[1])
        add(args[1], -(ybar * y / m[1]))
    elif op is Op.Inv:
        add(args[0], -(ybar * y * y))
    elif op is Op.Sqrt:
        add(args[0], ybar / (2 * y))
    elif op is Op.Max:                # std::max(a, b) = (a < b) ? b : a
        c = B.bind(m[0].lt(m[1]))
        add(args[0], B.select(c, zero, ybar))
        add(args[1], B.select(c, ybar, zero))
    elif op is Op.Min:                # std::min(a, b) = (b < a) ? b : a
        c = B.bind(m[1].lt(m[0]))
        add(args[0], B.select(c, zero, ybar))
        add(args[1], B.select(c, ybar, zero))
    elif op is Op.Abs:
        add(args[0], B.select(m[0].lt(0), -ybar, ybar))
    elif op is Op.Select:
        add(args[1], B.select(m[0], ybar, zero))
        add(args[2], B.select(m[0], zero, ybar))
    elif op is Op.Cast:
        src = g.nodes[args[0]].dyn.ctype
        add(args[0], B.cast(ybar, src))
    elif op is Op.Identity:
        add(args[0], ybar)
    elif op is Op.Exp:
        add(args[0], ybar * y)
    elif op is Op.Log:
        add(args[0], ybar / m[0])
    elif op is Op.Sin:
        add(args[0], ybar * B.cos(m[0]))
    elif op is Op.Cos:
        add(args[0], -(ybar * B.sin(m[0])))
    elif op in (Op.EQ, Op.NE, Op.LT, Op.LE, Op.GT, Op.GE, Op.And, Op.Or, Op.Not):
        return
    else:
        raise NotImplementedError("no adjoint rule for %s" % op.value)
