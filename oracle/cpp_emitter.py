"""ORACLE -- test infrastructure only; never imported by the product path.

CPU C++ emitter: a restatement of the reference's ``Native.CPlusPlus``
generator, ``Language/Paraiso/Generator/PlanTrans.hs:50-720``.

It consumes the same ``Plan`` as the HIP backend and emits the class the
reference would: one ``std::vector`` per Array storage, one member function
per subkernel with the per-cell SASA loop (``loopMaker``, PlanTrans.hs:406-596)
under ``#pragma omp parallel for``, the Arith -> C mapping of ``rhsArith``
(PlanTrans.hs:670-711), the sequential ``om_reduce_*`` library
(PlanTrans.hs:719) and deep-copy stores after all subkernels
(PlanTrans.hs:225-258).  A small ``extern "C"`` block is appended so tests can
drive the class through ctypes.

Deliberate, documented differences:
* ``abs_mode="fabs"`` (default) emits ``std::fabs`` for floating ``Abs``; the
  reference emits unqualified ``abs``, which g++ resolves to ``int abs(int)``
  with the generated includes (truncating).  ``abs_mode="reference"`` keeps it.
* scalar storages are zero-initialised (the reference leaves them
  uninitialised, so reading one before a store is undefined there).
"""
from __future__ import annotations

import hashlib
import os
import subprocess
import sys
from typing import Dict, List, Tuple

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from paraiso_amd import annotation as A                      # noqa: E402
from paraiso_amd.om.arith import Op                          # noqa: E402
from paraiso_amd.om.dyn import CType, Realm                  # noqa: E402
from paraiso_amd.om.graph import (Arith, Broadcast, Imm, Load, LoadIndex,  # noqa: E402
                                  LoadSize, Reduce, Shift, Store)
from paraiso_amd.om.literal import c_literal                 # noqa: E402
from paraiso_amd.generator.plan import Plan, SubKernel       # noqa: E402

LIBRARY = r"""
template <class T> T om_broadcast (const T& x) {
  return x;
}
template <class T> T om_reduce_sum (const std::vector<T> &xs) {
  T ret = 0;
  for (int i = 0; i < (int)xs.size(); ++i) ret+=xs[i];
  return ret;
}
template <class T> T om_reduce_min (const std::vector<T> &xs) {
  T ret = xs[0];
  for (int i = 1; i < (int)xs.size(); ++i) ret=std::min(ret,xs[i]);
  return ret;
}
template <class T> T om_reduce_max (const std::vector<T> &xs) {
  T ret = xs[0];
  for (int i = 1; i < (int)xs.size(); ++i) ret=std::max(ret,xs[i]);
  return ret;
}
"""


def _p(x: str) -> str:
    return "(" + x + ")"


def cursor_name(idx: int, cursor: Tuple[int, ...]) -> str:
    # PlanTrans.hs:630-644
    return "a%d_%s" % (idx, "_".join(str(c).replace("-", "m") for c in cursor))


def _type(ct: CType) -> str:
    return ct.c_name


def rhs_arith(op: Op, args: List[str], cast_to, abs_mode: str, arg_types) -> str:
    """PlanTrans.hs:670-711 (CPlusPlus branch)."""
    infix = {Op.Add: "+", Op.Sub: "-", Op.Mul: "*", Op.Div: "/", Op.Mod: "%",
             Op.And: "&&", Op.Or: "||", Op.EQ: "==", Op.NE: "!=", Op.LT: "<",
             Op.LE: "<=", Op.GT: ">", Op.GE: ">="}
    if op is Op.Identity:
        return args[0]
    if op in infix:
        return "%s %s %s" % (_p(args[0]), infix[op], _p(args[1]))
    if op is Op.Neg:
        return "-" + _p(args[0])
    if op is Op.Inv:
        return "1/" + _p(args[0])
    if op is Op.Not:
        return "!" + _p(args[0])
    if op is Op.Select:
        return "%s ? %s : %s" % (_p(args[0]), _p(args[1]), _p(args[2]))
    if op is Op.Max:
        return "std::max(%s, %s)" % (args[0], args[1])
    if op is Op.Min:
        return "std::min(%s, %s)" % (args[0], args[1])
    if op is Op.Abs:
        if abs_mode == "fabs" and arg_types[0].is_float:
            return "std::fabs(%s)" % args[0]
        return "abs(%s)" % args[0]
    if op in (Op.Sqrt, Op.Exp, Op.Log, Op.Sin, Op.Cos, Op.Tan, Op.Asin, Op.Acos, Op.Atan):
        return "%s(%s)" % (op.value.lower(), args[0])
    if op is Op.Atan2:
        return "atan2(%s, %s)" % (args[0], args[1])
    if op is Op.Cast:
        return "(%s)(%s)" % (cast_to.c_name, args[0])
    return "%s(%s)" % (op.value.lower(), ", ".join(args))


class _SubEmitter:
    def __init__(self, plan: Plan, sub: SubKernel, abs_mode: str):
        self.plan = plan
        self.sub = sub
        self.abs_mode = abs_mode
        self.g = plan.om.kernels[sub.kernel_idx].graph
        self.setup = plan.setup
        self.inputs = set(sub.inputs)
        self.outputs = set(sub.outputs)
        self.all = self.inputs | self.outputs | set(sub.calcs)
        nd = self.setup.ndim
        self.mem = list(plan.memory_size())
        self.bnd = []
        for k in range(nd):
            if self.setup.boundary[k] is A.Condition.Open:
                self.bnd.append(self.mem[k] - sub.lower_boundary[k] - sub.upper_boundary[k])
            else:
                self.bnd.append(self.mem[k])
        self.lm = list(plan.lower_margin)

    # --- codecs (PlanTrans.hs:440-484)
    def codec_mod(self):
        out = []
        nd = len(self.bnd)
        for k in range(nd):
            if k == 0:
                x = "i"
            else:
                prod = 1
                for j in range(k):
                    prod *= self.bnd[j]
                x = "%s / %s" % (_p("i"), _p(str(prod)))
            if k != nd - 1:
                x = "%s %% %s" % (_p(x), _p(str(self.bnd[k])))
            out.append(x)
        return out

    def codec_mod_add(self):
        return ["%s + %s" % (_p(x), _p(str(self.lm[k]))) for k, x in enumerate(self.codec_mod())]

    def codec_addr(self):
        if self.mem == self.bnd:
            return "i"
        terms = []
        for k, x in enumerate(self.codec_mod_add()):
            prod = 1
            for j in range(k):
                prod *= self.mem[j]
            terms.append("%s * %s" % (_p(x), _p(str(prod))))
        return _fold_plus(terms)

    def protect(self, k, x):
        if self.setup.boundary[k] is A.Condition.Open:
            return x
        n = str(self.mem[k])
        return "%s %% %s" % (_p("%s + %s" % (_p(x), _p(n))), _p(n))

    def codec_load_index(self, cursor):
        out = []
        for k, x in enumerate(self.codec_mod()):
            off = self.lm[k] - self.sub.lower_boundary[k] - cursor[k]
            out.append(self.protect(k, "%s - %s" % (_p(x), _p(str(off)))))
        return out

    def codec_cursor(self, cursor):
        if all(b is A.Condition.Open for b in self.setup.boundary):
            shift = 0
            for k, c in enumerate(cursor):
                prod = 1
                for j in range(k):
                    prod *= self.mem[j]
                shift += c * prod
            return "%s + %s" % (_p("addr_origin"), _p(str(shift)))
        terms = []
        for k, x in enumerate(self.codec_mod_add()):
            prod = 1
            for j in range(k):
                prod *= self.mem[j]
            terms.append("%s * %s" % (_p(str(prod)),
                                      _p(self.protect(k, "%s + %s" % (_p(x), _p(str(cursor[k])))))))
        return _fold_plus(terms)

    # --- SASA request analysis (PlanTrans.hs:526-570)
    def inst_of(self, idx):
        return self.g.inst_of_value(idx)

    def rhs_and_request(self, idx, cursor):
        inst_id, inst = self.inst_of(idx)
        prepre = self.g.pre(inst_id)
        nd = self.g.nodes[idx]
        if idx in self.inputs:
            if self.sub.realm is Realm.Array:
                if nd.dyn.realm is Realm.Array:
                    return "%s[%s]" % (_p("a%d" % idx), self.codec_cursor(cursor)), []
                return "a%d" % idx, []
            return "a%d" % idx, []
        if isinstance(inst, Imm):
            return c_literal(inst.value, inst.ctype), []
        if isinstance(inst, Arith):
            args = [cursor_name(p, cursor) for p in prepre]
            types = [self.g.nodes[p].dyn.ctype for p in prepre]
            return (rhs_arith(inst.op, args, inst.cast_to, self.abs_mode, types),
                    [(p, cursor) for p in prepre])
        if isinstance(inst, Shift):
            c2 = tuple(c - v for c, v in zip(cursor, inst.vec))
            return cursor_name(prepre[0], c2), [(prepre[0], c2)]
        if isinstance(inst, LoadIndex):
            return self.codec_load_index(cursor)[inst.axis], []
        if isinstance(inst, LoadSize):
            return str(self.setup.local_size[inst.axis]), []
        if isinstance(inst, Reduce):
            return "om_reduce_%s(%s)" % (inst.op.value.lower(), "a%d" % prepre[0]), []
        if isinstance(inst, Broadcast):
            return "om_broadcast(%s)" % ("a%d" % prepre[0]), []
        # the reference's fallback for any other instruction (PlanTrans.hs:570)
        return "42 /* TODO : %s */" % inst, []

    def body(self) -> List[str]:
        vals = sorted(i for i in self.all if self.g.nodes[i].is_value)
        zero = tuple(0 for _ in range(self.setup.ndim))
        cursors: Dict[int, set] = {}
        for idx in sorted(vals, reverse=True):
            if idx in self.outputs:
                cursors[idx] = {zero}
                continue
            req = set()
            for jdx in vals:
                if jdx <= idx:
                    continue
                for cur in cursors[jdx]:
                    _, rq = self.rhs_and_request(jdx, cur)
                    for n, c in rq:
                        if n == idx:
                            req.add(c)
            cursors[idx] = req
        lines = []
        for idx in vals:
            ct = self.g.nodes[idx].dyn.ctype
            for cur in sorted(cursors[idx]):
                expr, _ = self.rhs_and_request(idx, cur)
                if idx in self.outputs:
                    if self.sub.realm is Realm.Array:
                        lines.append("%s = %s;" % (_p("%s[addr_origin]" % _p("a%d" % idx)), _p(expr)))
                    else:
                        lines.append("%s = %s;" % (_p("a%d" % idx), _p(expr)))
                else:
                    lines.append("%s %s = %s;" % (_type(ct), cursor_name(idx, cur), expr))
        return lines

    def emit(self, cls: str) -> Tuple[str, str]:
        args = []
        for idx in self.sub.inputs:
            args.append("const %s & a%d" % (_storage_type(self.g.nodes[idx].dyn), idx))
        for idx in self.sub.outputs:
            args.append("%s & a%d" % (_storage_type(self.g.nodes[idx].dyn), idx))
        sig = "void %s (%s)" % (self.sub.name, ", ".join(args))
        out = ["void %s::%s (%s) {" % (cls, self.sub.name, ", ".join(args))]
        if self.sub.realm is Realm.Scalar:
            out.extend(self.body())
        else:
            n = 1
            for b in self.bnd:
                n *= b
            out.append("#pragma omp parallel for")
            out.append("for (int i = 0; i < %d; i += 1) {" % n)
            out.append("int addr_origin = %s;" % self.codec_addr())
            out.extend(self.body())
            out.append("}")
        out.append("}")
        return sig, "\n".join(out)


def _fold_plus(terms):
    acc = terms[0]
    for t in terms[1:]:
        acc = "%s + %s" % (_p(acc), _p(t))
    return acc


def _storage_type(dyn) -> str:
    if dyn.realm is Realm.Array:
        return "std::vector<%s>" % dyn.ctype.c_name
    return dyn.ctype.c_name


def emit(plan: Plan, abs_mode: str = "fabs") -> str:
    cls = plan.name
    om = plan.om
    lines = ["#include <algorithm>", "#include <cmath>", "#include <vector>", "#include <cstddef>",
             LIBRARY,
             "/*\nlowerMargin = %s\nupperMargin = %s\n*/" % (plan.lower_margin, plan.upper_margin)]
    subs = [_SubEmitter(plan, s, abs_mode) for s in plan.subkernels]
    emitted = [s.emit(cls) for s in subs]
    mem_total = 1
    for m in plan.memory_size():
        mem_total *= m
    loc_total = 1
    for m in plan.setup.local_size:
        loc_total *= m
    lines.append("class %s {" % cls)
    lines.append("public:")
    for st in plan.storages:
        lines.append("  %s %s;" % (_storage_type(st.dyn), st.name))
    inits = []
    for st in plan.storages:
        if st.dyn.realm is Realm.Array:
            inits.append("%s(om_memory_size())" % st.name)
        else:
            inits.append("%s(0)" % st.name)
    lines.append("  %s () : %s {}" % (cls, ", ".join(inits)))
    lines.append("  int om_size () { return %d; }" % loc_total)
    lines.append("  int om_memory_size () { return %d; }" % mem_total)
    for k in range(plan.setup.ndim):
        lines.append("  int om_size_%d () { return %d; }" % (k, plan.setup.local_size[k]))
        lines.append("  int om_memory_size_%d () { return %d; }" % (k, plan.memory_size()[k]))
        lines.append("  int om_lower_margin_%d () { return %d; }" % (k, plan.lower_margin[k]))
        lines.append("  int om_upper_margin_%d () { return %d; }" % (k, plan.upper_margin[k]))
    for sig, _ in emitted:
        lines.append("  " + sig + ";")
    for k in om.kernels:
        lines.append("  void %s ();" % k.name)
    lines.append("};")
    for _, body in emitted:
        lines.append(body)
    # kernels: subkernel calls in group order, then deep-copy stores (PlanTrans.hs:225-276)
    for ki, k in enumerate(om.kernels):
        body = []
        g = k.graph
        for s in call_order(plan, ki):
            argnames = [_find_var(plan, ki, g, idx) for idx in s.inputs + s.outputs]
            body.append("%s(%s);" % (s.name, ", ".join(argnames)))
        for i, nd in enumerate(g.nodes):
            if isinstance(nd.inst, Store):
                pre = g.pre(i)[0]
                st = plan.storage_of_static(nd.inst.static)
                ma = plan.storage_of_manifest(ki, pre)
                body.append("%s = %s;" % (st.name, ma.name))
        lines.append("void %s::%s () {\n%s\n}" % (cls, k.name, "\n".join(body)))
    # C ABI for ctypes-driven tests
    lines.append('extern "C" {')
    lines.append("void* oracle_create() { return new %s(); }" % cls)
    lines.append("void oracle_destroy(void* h) { delete (%s*)h; }" % cls)
    lines.append("void oracle_run(void* h, int k) {")
    lines.append("  %s* m = (%s*)h;" % (cls, cls))
    lines.append("  switch (k) {")
    for ki, k in enumerate(om.kernels):
        lines.append("    case %d: m->%s(); break;" % (ki, k.name))
    lines.append("  }")
    lines.append("}")
    lines.append("void oracle_run_n(void* h, int k, int n) { for (int t = 0; t < n; ++t) oracle_run(h, k); }")
    lines.append("void* oracle_static_ptr(void* h, int s) {")
    lines.append("  %s* m = (%s*)h;" % (cls, cls))
    lines.append("  switch (s) {")
    for si, sv in enumerate(om.setup.static_values):
        st = plan.storage_of_static(si)
        if sv.dyn.realm is Realm.Array:
            lines.append("    case %d: return (void*)m->%s.data();" % (si, st.name))
        else:
            lines.append("    case %d: return (void*)&m->%s;" % (si, st.name))
    lines.append("  }")
    lines.append("  return nullptr;")
    lines.append("}")
    lines.append("int oracle_memory_size() { return %d; }" % mem_total)
    lines.append("}")
    return "\n".join(lines) + "\n"


def call_order(plan: Plan, ki: int) -> List[SubKernel]:
    """Subkernel calls of kernel ``ki``: write-group order, except that a subkernel
    never runs before one whose outputs it reads.

    The reference calls subkernels in group order (PlanTrans.hs:233-241), but
    its grouping (DependencyAnalysis.hs:111-117) can put a Manifest node in an
    earlier group than a Manifest node it depends on: in the current Hydro,
    ``broadcast (cast (loadSize axis))`` lands in group 0 while the cast it reads
    is in group 1, so the reference's first ``proceed`` reads an uninitialised
    scalar and later calls read the previous call's value.  Ordering by
    dependency gives that steady-state behaviour from the first call on; it is
    the identity order whenever the grouping is consistent.
    """
    subs = [s for s in plan.subkernels if s.kernel_idx == ki]
    produced = {}
    for s in subs:
        for o in s.outputs:
            produced[o] = s.group
    done, order = set(), []
    while len(order) < len(subs):
        progressed = False
        for s in subs:
            if s.group in done:
                continue
            deps = set(produced[i] for i in s.inputs if i in produced and produced[i] != s.group)
            if deps <= done:
                order.append(s)
                done.add(s.group)
                progressed = True
                break
        if not progressed:
            raise ValueError("cyclic subkernel dependencies in kernel %d" % ki)
    return order


def _find_var(plan: Plan, ki: int, g, idx: int) -> str:
    """``findVar`` (PlanTrans.hs:259-276): a Manifest storage, or the static a Load reads."""
    for st in plan.storages:
        if st.idx.kind == "manifest" and st.idx.a == ki and st.idx.b == idx:
            return st.name
    for p in g.pre(idx):
        inst = g.nodes[p].inst
        if isinstance(inst, Load):
            return plan.storage_of_static(inst.static).name
    raise KeyError("no storage for node %d" % idx)


BUILD_DIR = os.path.join(HERE, "_build")


def build(plan: Plan, abs_mode: str = "fabs", opt: str = "-O2", openmp: bool = True) -> str:
    """Emit, compile with g++ (no FP contraction) and return the .so path."""
    src = emit(plan, abs_mode)
    flags = [opt, "-ffp-contract=off", "-fPIC", "-shared", "-std=c++17", "-w"]
    if openmp:
        flags.append("-fopenmp")
    key = hashlib.sha1((src + " ".join(flags)).encode()).hexdigest()[:16]
    os.makedirs(BUILD_DIR, exist_ok=True)
    so = os.path.join(BUILD_DIR, "%s_%s.so" % (plan.name, key))
    if os.path.exists(so):
        return so
    cpp = so[:-3] + ".cpp"
    with open(cpp, "w") as f:
        f.write(src)
    tmp = so + ".tmp%d" % os.getpid()
    subprocess.run(["g++"] + flags + [cpp, "-o", tmp], check=True)
    os.replace(tmp, so)
    return so
