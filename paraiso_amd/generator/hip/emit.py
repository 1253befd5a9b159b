"""HIP source emitter for scheduled OM kernels (gfx950).

For every launch of every OM kernel it writes one ``__global__`` function built
from the hand-written templates in ``csrc/om_device.h``:

1. tile coordinates from the workgroup id, remapped so that each of the
   eight XCDs owns a contiguous band of tiles (neighbouring tiles share halo
   data in that XCD's L2),
2. uniform scalars: scalar statics and slots read once per workgroup, and
   the scalar arithmetic of level 0 (the reference's Scalar subkernels,
   ``PlanTrans.hs:417``) evaluated redundantly in registers,
3. one cooperative loop per phase over the phase's box, with coalesced
   x-fastest addressing; LDS receives the materialised values,
4. the store phase over the tile: stores go to the ping-pong "next" buffers,
   Reduce arguments fold into per-thread accumulators (FP64 for floating
   sums),
5. the reduction epilogue: wave64 shuffle tree and one partial per workgroup;
   a one-workgroup ``_fin`` kernel, ordered after it on the stream, folds the
   partials in a fixed order, runs the scalar tail (``om_reduce_*`` /
   ``om_broadcast``, ``PlanTrans.hs:566-569``) and, in the kernel's last launch,
   commits the scalar stores.  Stream order replaces a device-wide fence per
   workgroup, which on a multi-XCD part would write back L2 thousands of times.

Per-point expressions follow the reference's SASA naming (``a<node>_<cursor>``,
``PlanTrans.hs:625-644``) as ``n<value>_<cursor>``.
"""
from __future__ import annotations

import re
from fractions import Fraction
from typing import Dict, List, Optional, Set, Tuple

import numpy as np

from ...om.arith import Op, ReduceOp
from ...om.dyn import CType, Realm, fraction_to_float32
from ...om.literal import c_literal
from .ir import LKernel
from .schedule import OUT, ArrayLaunch, KernelSchedule, ring_depth

NT = 256
PR_CTYPE = {CType.Bool: "PR_BOOL", CType.Int: "PR_INT", CType.Float: "PR_FLOAT",
            CType.Double: "PR_DOUBLE"}


def _cname(c: Tuple[int, ...]) -> str:
    return "_".join(("m%d" % -x) if x < 0 else str(x) for x in c)


def _p(x: str) -> str:
    return "(" + x + ")"


def acc_type(op: ReduceOp, ct: CType) -> CType:
    if op is ReduceOp.Sum and ct.is_float:
        return CType.Double
    if ct is CType.Bool:
        return CType.Int
    return ct


def acc_init(op: ReduceOp, ct: CType) -> str:
    if op is ReduceOp.Sum:
        return "0" if ct is CType.Int else "0.0"
    big = {CType.Int: ("2147483647", "(-2147483647-1)"),
           CType.Float: ("INFINITY", "-INFINITY"),
           CType.Double: ("((double)INFINITY)", "(-(double)INFINITY)")}[ct]
    return big[0] if op is ReduceOp.Min else big[1]


def acc_op(op: ReduceOp, ct: CType) -> str:
    if op is ReduceOp.Sum:
        return "OmSumWrap()" if ct is CType.Int else "OmSum()"
    return "OmMin()" if op is ReduceOp.Min else "OmMax()"


def _split_const(x: Fraction) -> Tuple[np.float32, np.float32]:
    """binary32 pair (hi, lo) with hi the nearest to x and lo the nearest to x - hi."""
    hi = fraction_to_float32(x)
    return hi, fraction_to_float32(x - Fraction(float(hi)))


def _const_product(x: str, c: Fraction, compensate=None) -> str:
    """x * c in FP32 for an exact rational c: a plain product when c is a binary32 value,
    else fma(x, hi, x * lo), which rounds x * c correctly but for near-ties.  A product by
    the rounded constant instead (0.4f for 2/5, 1.4f for 7/5) is off by the constant's own
    rounding error, the same sign in every cell and every step.  ``compensate`` (a set of
    constants, or None for all) limits the compensated form to the constants it holds; the
    others are multiplied by their nearest binary32 value."""
    hi, lo = _split_const(c)
    if lo == 0 or (compensate is not None and c not in compensate):
        return "%s * %s" % (_p(x), c_literal(hi, CType.Float))
    return "__builtin_fmaf(%s, %s, %s * %s)" % (x, c_literal(hi, CType.Float), _p(x),
                                               c_literal(lo, CType.Float))


def inverse_pairs(k: LKernel) -> Set[Fraction]:
    """The constants a kernel multiplies by (a product, or a division by the reciprocal)
    whose reciprocal it also multiplies by: c and 1/c convert a value forth and back (the
    pressure and the energy of a gas through gamma - 1), and unless x * c is correctly
    rounded the round trip drifts by the same factor every step.  Other constants
    (gamma in a sound speed, a wave-speed estimate) enter no round trip and are multiplied
    by their nearest binary32 value."""
    mult = set()
    for n in k.nodes:
        if n.op != "arith" or n.payload[0] not in (Op.Mul, Op.Div) or n.ctype is not CType.Float:
            continue
        a, b = (k.exact.get(x) for x in n.args[:2])
        if n.payload[0] is Op.Mul:
            mult.update(c for c in (a, b) if c is not None and c != 0)
        elif b is not None and b != 0:
            mult.add(1 / b)
    return set(c for c in mult if 1 / c in mult)


def device_arith(op: Op, args: List[str], cast_to, arg_types: List[CType], out: CType,
                 fast_fp32: bool = False, consts=None, compensate=None) -> str:
    """C expression of one operation.  ``consts`` holds, per argument, the exact rational
    value of a constant argument (exact-folding builds) or None; ``compensate`` as for
    _const_product."""
    consts = consts or [None] * len(args)
    if fast_fp32 and out is CType.Float and op in (Op.Mul, Op.Div) and any(
            c is not None for c in consts):
        if op is Op.Mul and consts[1] is not None and consts[0] is None:
            return _const_product(args[0], consts[1], compensate)
        if op is Op.Mul and consts[0] is not None and consts[1] is None:
            return _const_product(args[1], consts[0], compensate)
        if op is Op.Div and consts[1] is not None and consts[0] is None and consts[1] != 0:
            return _const_product(args[0], 1 / consts[1], compensate)
        if op is Op.Div and consts[0] is not None and consts[1] is None:
            return _const_product("__builtin_amdgcn_rcpf(%s)" % args[1], consts[0], compensate)
    infix = {Op.Add: "+", Op.Sub: "-", Op.Mul: "*", Op.Div: "/", Op.Mod: "%",
             Op.And: "&&", Op.Or: "||", Op.EQ: "==", Op.NE: "!=", Op.LT: "<",
             Op.LE: "<=", Op.GT: ">", Op.GE: ">="}
    t0 = arg_types[0] if arg_types else out
    if op is Op.Identity:
        return args[0]
    if op is Op.Div and t0 is CType.Float and fast_fp32:
        # v_rcp_f32 (1 ulp) and a multiply instead of the scaled IEEE division
        return "%s * __builtin_amdgcn_rcpf(%s)" % (_p(args[0]), args[1])
    if op is Op.Mod and out.is_float:
        raise NotImplementedError("Mod of floating values: the reference prints C's %, "
                                  "which takes integers only (PlanTrans.hs:678)")
    if op in infix:
        return "%s %s %s" % (_p(args[0]), infix[op], _p(args[1]))
    if op is Op.Neg:
        return "-" + _p(args[0])
    if op is Op.Inv:
        if t0 is CType.Float and fast_fp32:
            return "__builtin_amdgcn_rcpf(%s)" % args[0]
        return "1/" + _p(args[0])
    if op is Op.Not:
        return "!" + _p(args[0])
    if op is Op.Select:
        return "%s ? %s : %s" % (_p(args[0]), _p(args[1]), _p(args[2]))
    if op is Op.Max:
        return "om_max<%s>(%s, %s)" % (out.c_name, args[0], args[1])
    if op is Op.Min:
        return "om_min<%s>(%s, %s)" % (out.c_name, args[0], args[1])
    if op is Op.Abs:
        if t0 is CType.Float:
            return "fabsf(%s)" % args[0]
        if t0 is CType.Double:
            return "fabs(%s)" % args[0]
        return "abs(%s)" % args[0]
    if op is Op.Signum:
        return "(%s)(((%s) > 0) - ((%s) < 0))" % (out.c_name, args[0], args[0])
    if op is Op.Sqrt:
        if t0 is CType.Float and fast_fp32:
            return "__builtin_amdgcn_sqrtf(%s)" % args[0]
        return ("sqrtf(%s)" if t0 is CType.Float else "sqrt(%s)") % args[0]
    if op in (Op.Exp, Op.Log, Op.Sin, Op.Cos, Op.Tan, Op.Asin, Op.Acos, Op.Atan):
        f = op.value.lower()
        if t0 is CType.Float:
            # the reference calls the double libm function on float operands
            return "(float)%s((double)(%s))" % (f, args[0])
        return "%s(%s)" % (f, args[0])
    if op in (Op.Atan2, Op.Pow):
        # the reference prints atan2(x, y) and, through its fallback, pow(x, y)
        # (PlanTrans.hs:702,706): the double functions, also on float operands
        f = op.value.lower()
        if t0 is CType.Float:
            return "(float)%s((double)(%s), (double)(%s))" % (f, args[0], args[1])
        return "%s(%s, %s)" % (f, args[0], args[1])
    if op is Op.Cast:
        return "(%s)(%s)" % (cast_to.c_name, args[0])
    if op is Op.Mod and out.is_float:
        raise NotImplementedError("Mod of floating values: the reference prints C's %, "
                                  "which takes integers only (PlanTrans.hs:678)")
    raise NotImplementedError(
        "operator %s has no HIP lowering: the Builder never makes it, and the reference "
        "prints it as a call to an undefined function %s (PlanTrans.hs:706)"
        % (op.value, op.value.lower()))


class MachineEmitter:
    def __init__(self, name: str, statics, ndim: int, schedules: List[KernelSchedule],
                 tile: Tuple[int, ...], fast_fp32: bool = False, batched: bool = False,
                 threads: int = NT, open_axes: Tuple[bool, ...] = (), margins=None,
                 regions=None, march: int = 0, prefetch: bool = True, persist: bool = False,
                 forms: Optional[List[dict]] = None):
        """``forms``: per kernel, the launch form overriding the machine's (keys ``tile``,
        ``threads``, ``persist``, ``march``), or None."""
        self.name = name
        self.nt = threads
        # march > 0: a workgroup streams along the last axis through its tile, `march`
        # planes per step, each materialised value kept as a ring of planes in LDS
        self.march = march
        # with march: issue the global loads of step s + 1 into registers before step s
        # computes, so their latency overlaps the step's arithmetic
        self.prefetch = prefetch
        # register sets of streaming prefetch: the loads of step s + depth issue at step s
        # (two: the 3-D wave's launches 0.7-1.3% faster than with one, results unchanged)
        self.prefetch_depth = 2
        # block tiles: a resident grid of workgroups loops over the tiles (see emit_launch)
        self.persist = persist and not march
        self._m = None
        # Open (non-periodic) axes: grids carry lower/upper margins of ghost cells; a store or
        # reduce argument computed over region [lb, n - ub) holds 0 elsewhere
        self.open_axes = tuple(open_axes) + (False,) * (3 - len(open_axes))
        lm, um = margins if margins else ((0,) * ndim, (0,) * ndim)
        self.lm = tuple(lm) + (0,) * (3 - len(lm))
        self.um = tuple(um) + (0,) * (3 - len(um))
        self.regions = regions or {}     # (kernel, "store", static) / (kernel, "reduce", node)
        self._reach = [[0, 0] for _ in range(3)]   # offsets seen by _wrap in the current phase
        self._direct = False
        self._pf = None         # loop point whose prefetched registers stand for the loads
        self._comp = None       # constants whose products are compensated (None: all)
        self.fast_fp32 = fast_fp32
        self.batched = batched
        self.statics = statics           # list of Named
        self.ndim = ndim
        self.schedules = schedules
        self.tile = tuple(tile) + (1,) * (3 - len(tile))
        self._base = {"tile": self.tile, "threads": self.nt, "persist": self.persist,
                      "march": self.march}
        self.forms = forms or [None] * len(schedules)
        self.ns = len(statics)
        self.n_user = len(statics)
        # scratch arrays of cut launches become internal Array statics after the user's
        from ...om.dyn import DynValue
        from ...om.graph import Named
        from .schedule import SCRATCH_BASE
        self.statics = list(statics)
        self.scratch_map: Dict[Tuple[int, int], int] = {}
        for ki, sch in enumerate(schedules):
            for j, (v, ct) in enumerate(sch.scratch):
                self.scratch_map[(ki, SCRATCH_BASE + j)] = len(self.statics)
                self.statics.append(Named("__scratch_k%d_%d" % (ki, j), DynValue(Realm.Array, ct)))
        self.ns = len(self.statics)
        # slot numbering across kernels
        self.slot_index: Dict[Tuple[int, int], int] = {}
        for ki, sch in enumerate(schedules):
            for v in sch.slot_scalars:
                self.slot_index[(ki, v)] = len(self.slot_index)

    # ------------------------------------------------------------ helpers
    def _use_form(self, ki: int):
        """Tile, work-items, resident loop and streaming of kernel ki's launches."""
        f = dict(self._base)
        f.update(self.forms[ki] or {})
        self.tile = tuple(f["tile"]) + (1,) * (3 - len(f["tile"]))
        self.nt = int(f["threads"])
        self.march = int(f["march"])
        self.persist = bool(f["persist"]) and not self.march

    def fn_name(self, ki: int, launch: int) -> str:
        """Kernel symbol: machine, OM kernel and launch, so that profiles name them."""
        clean = lambda x: re.sub(r"[^A-Za-z0-9_]", "_", x)
        return "%s_%s_l%d" % (clean(self.name), clean(self.schedules[ki].kernel.name), launch)

    def sidx(self, ki, s) -> int:
        return self.scratch_map.get((ki, s), s)

    def is_scratch(self, s) -> bool:
        return s >= self.n_user

    def slot_off(self, ki, v) -> int:
        return 8 * (self.ns + self.slot_index[(ki, v)])

    def scalar_expr(self, k: LKernel, v: int) -> str:
        n = k.nodes[v]
        if n.op == "load":
            return "*(const %s*)(a.scal + %d)" % (n.ctype.c_name, 8 * n.payload)
        if n.op == "imm":
            return c_literal(n.payload, n.ctype)
        if n.op == "size":
            return "(%s)%s" % (n.ctype.c_name, self._local_size(n.payload))
        if n.op == "arith":
            op, cast_to = n.payload
            return device_arith(op, ["u%d" % x for x in n.args], cast_to,
                                [k.nodes[x].ctype for x in n.args], n.ctype, self.fast_fp32,
                                [k.exact.get(x) for x in n.args])
        raise NotImplementedError("scalar %s" % n.op)

    # ------------------------------------------------------------ per launch
    def emit_launch(self, ki: int, sch: KernelSchedule, la: ArrayLaunch) -> Tuple[str, str]:
        k = la.kernel_override or sch.kernel
        nodes = k.nodes
        nd = self.ndim
        T = self.tile
        fn = self.fn_name(ki, la.index)
        self._ki = ki
        self._comp = inverse_pairs(k) if self.fast_fp32 else None
        self._launch_reach = 0
        has_work = bool(la.targets)
        L = la.index
        # ---- argument struct
        loads = sorted(set(self.sidx(ki, nodes[v].payload) for v in la.lc if nodes[v].op == "load"
                           and nodes[v].realm is Realm.Array))
        stores = sorted(set(self.sidx(ki, s) for s, _ in la.stores))
        fields = []
        for s in loads:
            fields.append("const %s* __restrict__ ld%d;" % (self.statics[s].dyn.ctype.c_name, s))
        for s in stores:
            fields.append("%s* __restrict__ st%d;" % (self.statics[s].dyn.ctype.c_name, s))
        fields += ["char* __restrict__ scal;", "char* __restrict__ part;", "unsigned* counter;",
                   "int n0, n1, n2;", "int t0, t1, t2;", "unsigned ntiles;",
                   "unsigned m0, m1;  /* om_udiv reciprocals of t0 and t1 */",
                   "int wide;  /* some axis shorter than tile plus reach: om_wrap takes the "
                   "remainder */",
                   "unsigned nwg;   /* reduction partials written (workgroups that reduce) */"]
        struct = "struct Args_%s {\n  %s\n};\n" % (fn, "\n  ".join(fields))

        decls: List[str] = []
        if la.lds_bytes:
            decls.append("  __shared__ __align__(16) char smem[%d];" % la.lds_bytes)
        acc_types = set()
        for r in la.reduces:
            acc_types.add(acc_type(nodes[r].payload, nodes[nodes[r].args[0]].ctype))
        for at in sorted(acc_types, key=lambda c: c.value):
            decls.append("  __shared__ %s om_scr_%s[%d];" % (at.c_name, at.c_name,
                                                             self.nt // 64))
        coords = self._tile_coords("tile")
        body: List[str] = []
        w = body.append
        # uniform scalars
        for n in nodes:
            if n.realm is not Realm.Scalar:
                continue
            lvl = sch.level[n.id]
            if lvl == 0:
                w("  const %s u%d = %s;" % (n.ctype.c_name, n.id, self.scalar_expr(k, n.id)))
            elif lvl <= L:
                w("  const %s u%d = *(const %s*)(a.scal + %d);" % (
                    n.ctype.c_name, n.id, n.ctype.c_name, self.slot_off(ki, n.id)))
        # LDS views
        for v, b in sorted(la.lds.items()):
            w("  %s* __restrict__ s%d = (%s*)(smem + %d);" % (
                nodes[v].ctype.c_name, v, nodes[v].ctype.c_name, b.offset))
        # accumulators
        for r in la.reduces:
            a_ct = nodes[nodes[r].args[0]].ctype
            at = acc_type(nodes[r].payload, a_ct)
            w("  %s acc%d = %s;" % (at.c_name, r, acc_init(nodes[r].payload, at)))
        pre = list(body)             # tile-independent: scalars, LDS views, accumulators
        work: List[str] = []
        if has_work and self.march:
            self._emit_march(k, la, work.append)
        elif has_work:
            for ph in la.phases:
                self._emit_phase(k, la, ph, work.append)
        persist = has_work and self.persist
        if persist:
            loop = self._persistent_loop(k, la)
        # ---- epilogue: per-workgroup partials; the finish kernel does the rest
        # the finish kernel: reductions, scalars computed from them, and the commit of the
        # kernel's scalar stores (a launch with none of these has no finish kernel)
        needs_tail = bool(la.reduces) or bool(la.tail_scalars) or (la.commit and bool(sch.scalar_stores))

        def epilogue(index):
            e = []
            for j, r in enumerate(la.reduces):
                a_ct = nodes[nodes[r].args[0]].ctype
                at = acc_type(nodes[r].payload, a_ct)
                e.append("  { const %s br = om_block_reduce<%d>(acc%d, %s, om_scr_%s);" % (
                    at.c_name, self.nt, r, acc_op(nodes[r].payload, at), at.c_name))
                e.append("    if (threadIdx.x == 0) ((%s*)a.part)[%d * a.ntiles + %s] = br; }" % (
                    at.c_name, j, index))
            return e if has_work else []
        fin = []
        if needs_tail:
            f = fin.append
            for n in nodes:
                if n.realm is not Realm.Scalar:
                    continue
                lvl = sch.level[n.id]
                if lvl == 0:
                    f("  const %s u%d = %s;" % (n.ctype.c_name, n.id, self.scalar_expr(k, n.id)))
                elif lvl <= L:
                    f("  const %s u%d = *(const %s*)(a.scal + %d);" % (
                        n.ctype.c_name, n.id, n.ctype.c_name, self.slot_off(ki, n.id)))
            for j, r in enumerate(la.reduces):
                a_ct = nodes[nodes[r].args[0]].ctype
                at = acc_type(nodes[r].payload, a_ct)
                f("  const %s fr%d = om_partials_reduce<%d>((const %s*)a.part + %d * a.ntiles, "
                  "(int)a.nwg, (%s)%s, %s, om_scr_%s);" % (at.c_name, r, self.nt, at.c_name, j,
                                                         at.c_name,
                                              acc_init(nodes[r].payload, at),
                                              acc_op(nodes[r].payload, at), at.c_name))
            f("  if (threadIdx.x == 0) {")
            for v in la.tail_scalars:
                n = nodes[v]
                expr = "(%s)(fr%d)" % (n.ctype.c_name, v) if n.op == "reduce" else self.scalar_expr(k, v)
                f("    const %s u%d = %s;" % (n.ctype.c_name, v, expr))
                f("    *(%s*)(a.scal + %d) = u%d;" % (n.ctype.c_name, self.slot_off(ki, v), v))
            if la.commit:
                for s_, v in sch.scalar_stores:
                    f("    *(%s*)(a.scal + %d) = u%d;" % (nodes[v].ctype.c_name, 8 * s_, v))
            f("  }")
        fin_decls = ["  __shared__ %s om_scr_%s[%d];" % (at.c_name, at.c_name, self.nt // 64)
                     for at in sorted(acc_types, key=lambda c: c.value)]
        ks = [struct]
        core = "\n".join(coords + pre + work + epilogue("tile"))
        ks.append("extern \"C\" __global__ void __launch_bounds__(%d) %s(Args_%s a) {" % (self.nt, fn, fn))
        ks.extend(decls)
        if persist:
            ks.extend(pre)
            ks.extend(loop)
            ks.extend(epilogue("blockIdx.x"))
        else:
            ks.append("  unsigned tile = blockIdx.x;")
            ks.append("  if ((a.ntiles & 7u) == 0u) tile = (tile & 7u) * (a.ntiles >> 3) + (tile >> 3);")
            ks.append(core)
        ks.append("}")
        # batched form: the tiles of many grids in one launch (grid table in device memory)
        if self.batched:
            ks.append("extern \"C\" __global__ void __launch_bounds__(%d) %s_batched(const Args_%s* __restrict__ grids, "
                      "const int* __restrict__ tile_grid, const unsigned* __restrict__ tile_base, unsigned total) {"
                      % (self.nt, fn, fn))
            ks.extend(decls)
            ks.append("  unsigned b = blockIdx.x;")
            ks.append("  if ((total & 7u) == 0u) b = (b & 7u) * (total >> 3) + (b >> 3);")
            ks.append("  const int g = tile_grid[b];")
            ks.append("  const Args_%s a = grids[g];" % fn)
            ks.append("  const unsigned tile = b - tile_base[g];")
            ks.append(core)
            ks.append("}")
        if needs_tail:
            fcore = "\n".join(fin)
            ks.append("extern \"C\" __global__ void __launch_bounds__(%d) %s_fin(Args_%s a) {" % (self.nt, fn, fn))
            ks.extend(fin_decls)
            ks.append(fcore)
            ks.append("}")
            if self.batched:
                ks.append("extern \"C\" __global__ void __launch_bounds__(%d) %s_fin_batched(const Args_%s* __restrict__ grids) {"
                          % (self.nt, fn, fn))
                ks.extend(fin_decls)
                ks.append("  const Args_%s a = grids[blockIdx.x];" % fn)
                ks.append(fcore)
                ks.append("}")
        kernel_src = "\n".join(ks) + "\n"

        # ---- host launcher
        h = []
        h.append("static void fill_%s(const pr_launch_ctx* c, Args_%s& a) {" % (fn, fn))
        for s in loads:
            h.append("  a.ld%d = (const %s*)c->cur[%d];" % (s, self.statics[s].dyn.ctype.c_name, s))
        for s in stores:
            src = "cur" if self.is_scratch(s) else "next"
            h.append("  a.st%d = (%s*)c->%s[%d];" % (s, self.statics[s].dyn.ctype.c_name, src, s))
        h.append("  a.scal = c->scal; a.part = c->part; a.counter = c->counter;")
        h.append("  a.n0 = c->n[0]; a.n1 = c->n[1]; a.n2 = c->n[2];")
        h.append("  a.t0 = c->tiles[0]; a.t1 = c->tiles[1]; a.t2 = c->tiles[2]; a.ntiles = c->ntiles;")
        h.append("  a.m0 = om_udiv_magic((unsigned)c->tiles[0]); a.m1 = om_udiv_magic((unsigned)c->tiles[1]);")
        reach = self._max_reach(sch, la)
        h.append("  a.wide = (%s) ? 1 : 0;" % " || ".join(
            "c->n[%d] < %d" % (ax, self.tile[ax] + reach) for ax in range(self.ndim)))
        if persist:
            h.append("  a.nwg = om_resident_workgroups(c->ntiles, %d);" % self._per_cu(la))
        else:
            h.append("  a.nwg = c->ntiles;")
        h.append("}")
        h.append("static void pack_%s(const pr_launch_ctx* c, void* out) {" % fn)
        h.append("  Args_%s a; fill_%s(c, a); a.nwg = c->ntiles; memcpy(out, &a, sizeof a);" % (fn, fn))
        h.append("}")
        h.append("static int launch_batched_%s(const pr_batch_ctx* bc) {" % fn)
        if not self.batched:
            h.append("  (void)bc; return -1;")
        elif has_work:
            h.append("  hipLaunchKernelGGL(%s_batched, dim3(bc->total_tiles), dim3(%d), 0, (hipStream_t)bc->stream, "
                     "(const Args_%s*)bc->args, bc->tile_grid, bc->tile_base, bc->total_tiles);" % (fn, self.nt, fn))
        if needs_tail and self.batched:
            h.append("  hipLaunchKernelGGL(%s_fin_batched, dim3(bc->n_grids), dim3(%d), 0, (hipStream_t)bc->stream, "
                     "(const Args_%s*)bc->args);" % (fn, self.nt, fn))
        if self.batched:
            h.append("  return hipPeekAtLastError() == hipSuccess ? 0 : -1;")
        h.append("}")
        h.append("static int launch_%s(const pr_launch_ctx* c) {" % fn)
        h.append("  Args_%s a; fill_%s(c, a);" % (fn, fn))
        if has_work:
            h.append("  hipLaunchKernelGGL(%s, dim3(%s), dim3(%d), 0, (hipStream_t)c->stream, a);"
                     % (fn, "a.nwg" if persist else "c->ntiles", self.nt))
        if needs_tail:
            h.append("  hipLaunchKernelGGL(%s_fin, dim3(1), dim3(%d), 0, (hipStream_t)c->stream, a);"
                     % (fn, self.nt))
        h.append("  return hipPeekAtLastError() == hipSuccess ? 0 : -1;")
        h.append("}")
        return kernel_src, "\n".join(h) + "\n"

    def _max_reach(self, sch, la) -> int:
        """A bound on how far outside its tile a launch's coordinates go on any axis: the
        widest phase box plus the largest offset the launch wraps (recorded by _wrap while
        its phases were emitted)."""
        box = 0
        for ph in la.phases:
            for lo, hi in ph.box:
                box = max(box, -lo, hi)
        return box + max(sch.halo, self._launch_reach) + 1

    def _tile_coords(self, tile: str) -> List[str]:
        """Tile origin from the tile index: one division per axis after the first, by the
        host's reciprocal of the tile count (om_udiv), none for the last axis."""
        T = self.tile
        nd = self.ndim
        if nd == 1:
            return ["  const int X0 = (int)(%s) * %d, X1 = 0, X2 = 0;" % (tile, T[0]),
                    "  (void)X0; (void)X1; (void)X2;"]
        out = ["  const unsigned q0_ = om_udiv(%s, (unsigned)a.t0, a.m0);" % tile,
               "  const int X0 = (int)(%s - q0_ * (unsigned)a.t0) * %d;" % (tile, T[0])]
        if nd == 2:
            out.append("  const int X1 = (int)q0_ * %d, X2 = 0;" % T[1])
        else:
            out += ["  const unsigned q1_ = om_udiv(q0_, (unsigned)a.t1, a.m1);",
                    "  const int X1 = (int)(q0_ - q1_ * (unsigned)a.t1) * %d;" % T[1],
                    "  const int X2 = (int)q1_ * %d;" % T[2]]
        out.append("  (void)X0; (void)X1; (void)X2;")
        return out

    def _per_cu(self, la) -> int:
        """Workgroups of this launch one CU holds at once, as LDS and work-items allow."""
        per = 2048 // self.nt
        if la.lds_bytes:
            per = min(per, (160 * 1024) // la.lds_bytes)
        return max(1, per)

    def _persistent_loop(self, k, la) -> List[str]:
        """Single-grid body of a resident grid: workgroup b takes tiles b, b + nwg, ... (in
        the XCD-banded order of the one-tile-per-workgroup launch, nwg being a multiple of 8),
        accumulating its reductions across them.  When the first phase only loads fields,
        the loads of the next tile are issued into registers once the current tile's have
        been written to LDS, so they are in flight while the current tile computes."""
        out: List[str] = []
        w = out.append
        phases = la.phases
        pf = self.prefetch and len(phases) > 1 and phases[0].index == 0 and phases[0].nodes

        def remap(t):
            return ("const unsigned tile = ((a.ntiles & 7u) == 0u) ? ((%s) & 7u) * (a.ntiles >> 3) "
                    "+ ((%s) >> 3) : (%s);" % (t, t, t))
        if pf:
            self._emit_tile_loads(k, la, phases[0], w, declare=True, load=False)
            w("  if (blockIdx.x < a.ntiles) {")
            w("  " + remap("blockIdx.x"))
            out.extend(self._tile_coords("tile"))
            self._emit_tile_loads(k, la, phases[0], w, declare=False)
            w("  }")
        w("  for (unsigned tt = blockIdx.x; tt < a.ntiles; tt += a.nwg) {")
        w("  " + remap("tt"))
        out.extend(self._tile_coords("tile"))
        for pi, ph in enumerate(phases):
            if pf and pi == 0:
                self._emit_tile_writes(k, la, ph, w)
                w("  __syncthreads();")
                w("  if (tt + a.nwg < a.ntiles) {")
                w("  " + remap("tt + a.nwg"))
                out.extend(self._tile_coords("tile"))
                self._emit_tile_loads(k, la, ph, w, declare=False)
                w("  }")
                continue
            self._emit_phase(k, la, ph, w)
        w("  }")
        return out

    def _emit_tile_loads(self, k, la, ph, w, declare, load=True):
        """Global loads of a load-only phase of the tile at (X0, X1, X2) into registers
        pf<v>_<i>, one per loop point i of the thread."""
        T = self.tile
        box = tuple(ph.box) + ((0, 0),) * (3 - len(ph.box))
        W = [T[ax] + box[ax][1] - box[ax][0] for ax in range(3)]
        vol = W[0] * W[1] * W[2]

        def close(guarded):
            if guarded:
                w("    }")
            w("  }")
        self._emit_prefetch(k, la, ph, w, box, W, (vol + self.nt - 1) // self.nt, declare, load,
                            "", lambda i: self._tile_point(box, W, vol, i, w), close)

    def _emit_tile_writes(self, k, la, ph, w):
        T = self.tile
        box = tuple(ph.box) + ((0, 0),) * (3 - len(ph.box))
        W = [T[ax] + box[ax][1] - box[ax][0] for ax in range(3)]
        vol = W[0] * W[1] * W[2]
        n_it = (vol + self.nt - 1) // self.nt
        w("  // phase %d: values %s into LDS (loaded one tile ahead)"
          % (ph.index, ",".join("n%d" % v for v in ph.nodes)))
        computed = self._load_roots(k, la, ph) != sorted(ph.nodes)
        for i in range(n_it):
            if computed:
                guarded = self._tile_point(box, W, vol, i, w)
                self._pf = i
                for ln in self._ctx_body(k, la, list(ph.nodes), box, W, False):
                    w("    " + ln)
                self._pf = None
                for v in ph.nodes:
                    w("    s%d[p] = n%d_%s;" % (v, v, _cname((0,) * self.ndim)))
                if guarded:
                    w("    }")
                w("  }")
                continue
            w("  {")
            w("    const int p = threadIdx.x + %d;" % (i * self.nt))
            if (i + 1) * self.nt > vol:
                w("    if (p < %d)" % vol)
            w("    { %s }" % " ".join("s%d[p] = pf%d_%d;" % (v, v, i) for v in ph.nodes))
            w("  }")

    def _load_roots(self, k, la, ph) -> List[int]:
        """The fields a load phase reads from memory: its materialised loads, and the loads
        of the values it computes point by point from them."""
        nodes = k.nodes
        out = set(v for v in ph.nodes if nodes[v].op == "load")
        for u, ctxs in la.lc.items():
            if nodes[u].op == "load" and nodes[u].realm is Realm.Array and ph.index in ctxs:
                out.add(u)
        return sorted(out)

    def _tile_point(self, box, W, vol, i, w) -> bool:
        """Open the block of loop point p = threadIdx.x + i * NT of a block-tile phase."""
        w("  {")
        w("    const int p = threadIdx.x + %d;" % (i * self.nt))
        guarded = (i + 1) * self.nt > vol
        if guarded:
            w("    if (p < %d) {" % vol)
        w("    const int l0 = p %% %d;" % W[0])
        w("    const int l1 = (p / %d) %% %d;" % (W[0], W[1]))
        w("    const int l2 = p / %d;" % (W[0] * W[1]))
        w("    (void)l0; (void)l1; (void)l2;")
        for ax in range(3):
            w("    const int g%d = X%d + %d + l%d; (void)g%d;" % (ax, ax, box[ax][0], ax, ax))
        return guarded

    def _emit_march(self, k: LKernel, la: ArrayLaunch, w):
        """Stream along the last axis: at step s every phase computes the R planes at its
        leading edge, [s R + hi, s R + hi + R) of its box (planes below the box are
        skipped), then the store phase does planes [s R, s R + R).  A phase's ring holds
        hi - lo + R planes, so what later phases and the next steps read is still there."""
        R = self.march
        ma = self.ndim - 1
        Z = self.tile[ma]
        if Z % R:
            raise ValueError("the tile's last extent %d is not a multiple of march %d" % (Z, R))
        inner = la.phases[:-1]
        s0 = min([(ph.box[ma][0] - ph.box[ma][1]) // R for ph in inner] + [0])
        pf = self.prefetch and inner and inner[0].index == 0 and inner[0].nodes
        # with two register sets the loads of step s + 2 are issued at step s: two steps of
        # arithmetic to cover their latency (the loop is unrolled twice, the sets alternate)
        depth = 2 if pf and self.prefetch_depth >= 2 else 1
        N = Z // R
        w("  // march along axis %d: %d planes per step over %d planes" % (ma, R, Z))
        n_pf = 0
        if pf:
            ph = inner[0]
            self._m = {"R": R, "ma": ma, "lo": ph.box[ma][0], "hi": ph.box[ma][1]}
            n_pf = self._march_points(self._march_box(ph)[2])[2]
            for d in range(depth):
                self._emit_march_loads(k, la, ph, w, str(s0 + d), declare=True,
                                       base=d * n_pf)

        def step_body(base):
            for pi, ph in enumerate(inner):
                self._m = {"R": R, "ma": ma, "lo": ph.box[ma][0], "hi": ph.box[ma][1]}
                if pf and pi == 0:
                    self._emit_march_ring_writes(k, la, ph, w, base)
                    w("  __syncthreads();")
                    w("  if (s + %d < %d) {" % (depth, N))
                    self._emit_march_loads(k, la, ph, w, "(s + %d)" % depth, declare=False,
                                           base=base)
                    w("  }")
                    continue
                self._emit_phase(k, la, ph, w)
            self._m = {"R": R, "ma": ma, "lo": 0, "hi": 0}
            w("  if (s >= 0) {")
            self._emit_phase(k, la, la.phases[-1], w)
            w("  }")
        if depth == 1:
            w("  for (int s = %d; s < %d; ++s) {" % (s0, N))
            step_body(0)
            w("  }")
        else:
            w("  for (int s_ = %d; s_ < %d; s_ += 2) {" % (s0, N))
            w("  {")
            w("  const int s = s_;")
            step_body(0)
            w("  }")
            w("  if (s_ + 1 < %d) {" % N)
            w("  const int s = s_ + 1;")
            step_body(n_pf)
            w("  }")
            w("  }")
        self._m = None

    def _march_box(self, ph):
        T = self.tile
        box = tuple(ph.box) + ((0, 0),) * (3 - len(ph.box))
        W = [T[ax] + box[ax][1] - box[ax][0] for ax in range(3)]
        W[self._m["ma"]] = self._m["R"]
        return box, W, W[0] * W[1] * W[2]

    def _march_points(self, vol) -> Tuple[int, int, int, bool]:
        """(points of one plane, loop points per plane of a thread, loop points of a thread,
        by plane) of a march phase whose R leading planes hold ``vol`` points.  A thread walks
        the planes one at a time when that issues no more than 10% more wave64 passes than
        walking the R planes as one range; small planes (a 2-D row of a few waves) are walked
        as one range."""
        R = self._m["R"]
        pv = vol // R
        per = (pv + self.nt - 1) // self.nt
        by_plane = R * ((pv + 63) // 64) <= 1.1 * ((vol + 63) // 64)
        n = R * per if by_plane else (vol + self.nt - 1) // self.nt
        return pv, per, n, by_plane

    def _march_point(self, box, W, vol, j, step, w, is_out=False, width0=None):
        """Open the block of the thread's loop point j of march step `step`.  By plane: plane
        l of the box's leading R planes and point p = threadIdx.x + i * NT of that plane, (l,
        i) = divmod(j, points per plane); the plane coordinate jz is uniform over the
        workgroup and the in-plane coordinates do not depend on the step, so their index
        arithmetic is evaluated once per kernel rather than at every step.  As one range:
        point p = threadIdx.x + j * NT of the R planes, plane p / (points of a plane).  With
        ``width0`` below W[0], rows are walked W[0] points long and points at x >= width0
        are skipped.  The caller closes the block with _march_close."""
        m = self._m
        ma = m["ma"]
        pv, per, n, by_plane = self._march_points(vol)
        w("  {")
        conds = []
        if by_plane:
            l, i = divmod(j, per)
            w("    const int jz = %s * %d + %d;" % (step, m["R"], m["hi"] + l))
            w("    const int p = threadIdx.x + %d;" % (i * self.nt))
            if (i + 1) * self.nt > pv:
                conds.append("p < %d" % pv)
            lma = "%d" % l
        else:
            w("    const int p = threadIdx.x + %d;" % (j * self.nt))
            w("    const int jz = %s * %d + %d + p / %d;" % (step, m["R"], m["hi"], pv))
            if (j + 1) * self.nt > vol:
                conds.append("p < %d" % vol)
            lma = "p / %d" % pv
        if not is_out:
            conds.append("jz >= %d" % m["lo"])
        stride = 1
        for ax in range(3):
            if ax == ma:
                w("    const int l%d = %s;" % (ax, lma))
                continue
            if W[ax] == 1:
                w("    const int l%d = 0;" % ax)
                continue
            q = "p" if stride == 1 else "(p / %d)" % stride
            last = by_plane and stride * W[ax] >= pv
            w("    const int l%d = %s;" % (ax, q if last else "%s %% %d" % (q, W[ax])))
            stride *= W[ax]
        if width0 is not None and width0 < W[0]:
            conds.append("l0 < %d" % width0)
        w("    if (%s) {" % (" && ".join(conds) if conds else "true"))
        w("    (void)l0; (void)l1; (void)l2;")
        for ax in range(3):
            if ax == ma:
                w("    const int g%d = X%d + jz; (void)g%d;" % (ax, ax, ax))
            else:
                w("    const int g%d = X%d + %d + l%d; (void)g%d;" % (ax, ax, box[ax][0], ax, ax))
        return False

    def _march_close(self, guarded, w):
        w("    }")
        w("  }")

    def _emit_march_loads(self, k, la, ph, w, step, declare, base=0):
        """Global loads of a load-only phase for march step `step` into registers
        pf<v>_<base + i> (one per loop point i of the thread), which a later ring write
        stores."""
        box, W, vol = self._march_box(ph)
        self._emit_prefetch(k, la, ph, w, box, W, self._march_points(vol)[2],
                            declare, True, "for step %s" % step,
                            lambda j: self._march_point(box, W, vol, j, step, w),
                            lambda guarded: self._march_close(guarded, w), base)

    def _emit_prefetch(self, k, la, ph, w, box, W, n_it, declare, load, what, open_point,
                       close_point, base=0):
        """Loads of a load-only phase into registers pf<v>_<i>, i the thread's loop point:
        direct addressing for tiles whose reads stay inside the grid (a tile-uniform test),
        wrapped or clamped addressing for the others.  ``declare`` declares the registers
        first; with ``load`` False nothing else is emitted."""
        roots = self._load_roots(k, la, ph)
        if declare:
            for v in roots:
                tn = k.nodes[v].ctype.c_name
                w("  %s %s;" % (tn, ", ".join("pf%d_%d = (%s)0" % (v, base + i, tn)
                                              for i in range(n_it))))
        if not load:
            return
        self._reach = [[0, 0] for _ in range(3)]
        self._direct = False
        bodies = [self._ctx_body(k, la, roots, box, W, False)]
        conds = None
        # (a marching workgroup wraps every point: its plane coordinate is uniform and the
        # in-plane coordinates do not change from step to step, so the wrap is hoisted)
        if not self._m and any("om_wrap(" in ln or "om_clamp(" in ln for ln in bodies[0]):
            reach = self._reach
            self._direct = True
            bodies.insert(0, self._ctx_body(k, la, roots, box, W, False))
            self._direct = False
            conds = []
            T = self.tile
            for ax in range(self.ndim):
                lo = box[ax][0] + reach[ax][0]
                hi = T[ax] + box[ax][1] - 1 + reach[ax][1]
                conds.append("X%d + (%d) >= 0 && X%d + (%d) < a.n%d" % (ax, lo, ax, hi, ax))
        w("  // prefetch: loads of phase %d%s" % (ph.index, (" " + what) if what else ""))
        for bi, body in enumerate(bodies):
            if conds is not None:
                w("  if (%s) {" % " && ".join(conds) if bi == 0 else "  } else {")
            for i in range(n_it):
                guarded = open_point(i)
                for ln in body:
                    mt = re.match(r"const \w+ n(\d+)_0(?:_0)* = (.*)$", ln)
                    if not mt or int(mt.group(1)) not in roots:
                        raise ValueError("prefetch expects plain loads, got: %s" % ln)
                    w("      pf%s_%d = %s" % (mt.group(1), base + i, mt.group(2)))
                close_point(guarded)
        if conds is not None:
            w("  }")

    def _emit_march_ring_writes(self, k, la, ph, w, base=0):
        """Store the registers loaded for this step (pf<v>_<base + i>) into the value rings."""
        box, W, vol = self._march_box(ph)
        n_it = self._march_points(vol)[2]
        w("  // phase %d: values %s into their rings (loaded one step ahead)"
          % (ph.index, ",".join("n%d" % v for v in ph.nodes)))
        computed = self._load_roots(k, la, ph) != sorted(ph.nodes)
        zero = _cname((0,) * self.ndim)
        for i in range(n_it):
            guarded = self._march_point(box, W, vol, i, "s", w)
            if computed:
                # values computed point by point from the prefetched fields
                self._pf = base + i
                for ln in self._ctx_body(k, la, list(ph.nodes), box, W, False):
                    w("      " + ln)
                self._pf = None
            for v in ph.nodes:
                src = ("n%d_%s" % (v, zero)) if computed else ("pf%d_%d" % (v, base + i))
                w("      s%d[%s] = %s;" % (v, self._lds_self(la, v, W), src))
            self._march_close(guarded, w)

    def _ring(self, la, v):
        """(depth, plane size, bias) of value v's ring of planes."""
        ma = self.ndim - 1
        bb = tuple(la.lds[v].box) + ((0, 0),) * (3 - len(la.lds[v].box))
        depth = ring_depth(bb[ma][1] - bb[ma][0] + self.march)
        plane = 1
        for ax in range(3):
            if ax != ma:
                plane *= self.tile[ax] + bb[ax][1] - bb[ax][0]
        return depth, plane, depth * 64

    def _emit_phase(self, k: LKernel, la: ArrayLaunch, ph, w):
        nodes = k.nodes
        nd = self.ndim
        T = self.tile
        box = tuple(ph.box) + ((0, 0),) * (3 - len(ph.box))
        W = [T[ax] + box[ax][1] - box[ax][0] for ax in range(3)]
        if self._m:
            W[self._m["ma"]] = self._m["R"]          # one batch of planes per step
        vol = W[0] * W[1] * W[2]
        is_out = ph.nodes == [OUT]
        w("  // phase %d: %s over box %s" % (ph.index, "store" if is_out else
                                             "values " + ",".join("n%d" % v for v in ph.nodes),
                                             list(box[:nd])))
        roots = [t for t in la.targets if t not in la.mat] if is_out else list(ph.nodes)
        self._reach = [[0, 0] for _ in range(3)]
        self._direct = False
        wrapped = self._ctx_body(k, la, roots, box, W, is_out)
        reach = self._reach
        if any("om_wrap(" in ln for ln in wrapped):
            # tile-uniform test: every wrapped coordinate of this phase lies inside the grid,
            # so interior tiles index directly and only tiles on the border wrap
            self._direct = True
            direct = self._ctx_body(k, la, roots, box, W, is_out)
            self._direct = False
            conds = []
            for ax in range(nd):
                lo = box[ax][0] + reach[ax][0]
                hi = T[ax] + box[ax][1] - 1 + reach[ax][1]
                conds.append("X%d + (%d) >= 0 && X%d + (%d) < a.n%d" % (ax, lo, ax, hi, ax))
            w("  if (%s) {" % " && ".join(conds))
            self._emit_loop(k, la, ph, box, W, vol, is_out, roots, direct, w, "  ")
            w("  } else {")
            self._emit_loop(k, la, ph, box, W, vol, is_out, roots, wrapped, w, "  ")
            w("  }")
        else:
            self._emit_loop(k, la, ph, box, W, vol, is_out, roots, wrapped, w, "")
        w("  __syncthreads();")

    def _emit_loop(self, k, la, ph, box, W, vol, is_out, roots, lines, w, ind):
        nodes = k.nodes
        nd = self.ndim
        zero = tuple(0 for _ in range(nd))

        def body(ind):
            if is_out:
                conds = ["g%d < a.n%d" % (ax, ax) for ax in range(nd)]
                if self._m:
                    w(ind + "    if (%s) {" % " && ".join(conds))
                else:
                    w(ind + "    if (!(%s)) continue;" % " && ".join(conds))
                w(ind + "    const int addr0 = %s;" % self._addr(["g0", "g1", "g2"]))
            for ln in lines:
                w(ind + "    " + ln)
            if is_out:
                for s, v in la.stores:
                    val = self._val_name(k, la, v, zero, roots)
                    cond = self._region_cond((self._ki, "store", s))
                    if cond:
                        val = "(%s) ? %s : (%s)0" % (cond, val, nodes[v].ctype.c_name)
                    w(ind + "    a.st%d[addr0] = %s;" % (self.sidx(self._ki, s), val))
                for r in la.reduces:
                    arg = nodes[r].args[0]
                    at = acc_type(nodes[r].payload, nodes[arg].ctype)
                    val = "(%s)%s" % (at.c_name, self._val_name(k, la, arg, zero, roots))
                    cond = self._region_cond((self._ki, "reduce", r))
                    if cond:
                        val = "(%s) ? %s : (%s)0" % (cond, val, at.c_name)
                    w(ind + "    acc%d = %s(acc%d, %s);" % (r, acc_op(nodes[r].payload, at), r, val))
                if self._m:
                    w(ind + "    }")
            else:
                for v in ph.nodes:
                    w(ind + "    s%d[%s] = n%d_%s;" % (v, self._lds_self(la, v, W), v, _cname(zero)))

        if self._m:
            # a march phase: the thread's loop points, unrolled (see _march_point).  A phase
            # whose rows do not fill whole waves walks them as long as the widest LDS rows it
            # reads, so a wave's reads of those rows are consecutive words (no bank conflict
            # where a wave crosses from one row to the next)
            Wr = list(W)
            if W[0] % 64:
                T0 = self.tile[0]
                pitch = [T0 + b.box[0][1] - b.box[0][0] for v, b in la.lds.items()
                         if ph.index in la.lc.get(v, {}) or (is_out and OUT in la.lc.get(v, {}))]
                Wr[0] = max([W[0]] + pitch)
            volr = vol // W[0] * Wr[0]
            for j in range(self._march_points(volr)[2]):
                self._march_point(box, Wr, volr, j, "s", lambda x: w(ind + x), is_out, W[0])
                body(ind + "  ")
                self._march_close(False, lambda x: w(ind + x))
            return
        w(ind + "  for (int p = threadIdx.x; p < %d; p += %d) {" % (vol, self.nt))
        w(ind + "    const int l0 = p %% %d;" % W[0])
        w(ind + "    const int l1 = (p / %d) %% %d;" % (W[0], W[1]))
        w(ind + "    const int l2 = p / %d;" % (W[0] * W[1]))
        w(ind + "    (void)l0; (void)l1; (void)l2;")
        for ax in range(3):
            w(ind + "    const int g%d = X%d + %d + l%d; (void)g%d;" % (ax, ax, box[ax][0], ax, ax))
        body(ind)
        w(ind + "  }")

    def _lds_self(self, la, v, W) -> str:
        """Where a phase stores its own value at the loop point."""
        if not self._m:
            return "p"
        ma = self._m["ma"]
        depth, plane, bias = self._ring(la, v)
        terms, stride = [], 1
        for ax in range(3):
            if ax == ma:
                continue
            terms.append("l%d * %d" % (ax, stride))
            stride *= W[ax]
        return "((jz + %d) & %d) * %d + %s" % (bias, depth - 1, plane, " + ".join(terms))

    def _val_name(self, k, la, v, c, roots):
        return "n%d_%s" % (v, _cname(c))

    def _addr(self, g: List[str]) -> str:
        if self.ndim == 1:
            return g[0]
        if self.ndim == 2:
            return "%s * a.n0 + %s" % (g[1], g[0])
        # plane offset (uniform when the plane coordinate is) plus the in-plane offset
        return "%s * (a.n1 * a.n0) + (%s * a.n0 + %s)" % (g[2], g[1], g[0])

    def _ctx_body(self, k: LKernel, la: ArrayLaunch, roots: List[int], box, W, is_out) -> List[str]:
        nodes = k.nodes
        nd = self.ndim
        zero = tuple(0 for _ in range(nd))
        needed: Dict[int, Set[Tuple[int, ...]]] = {}
        rootset = set(roots)
        stack = [(r, zero) for r in roots]
        # in the store phase, materialised targets are read from LDS at offset 0
        if is_out:
            for t in la.targets:
                if t in la.mat:
                    stack.append((t, zero))
        while stack:
            v, c = stack.pop()
            s = needed.setdefault(v, set())
            if c in s:
                continue
            s.add(c)
            n = nodes[v]
            if v in la.mat and v not in rootset:
                continue
            if n.op == "shift":
                c2 = tuple(x - y for x, y in zip(c, n.payload))
                stack.append((n.args[0], c2))
            elif n.op == "arith":
                for a_ in n.args:
                    stack.append((a_, c))
        lines = []
        for v in sorted(needed):
            n = nodes[v]
            tn = n.ctype.c_name
            for c in sorted(needed[v]):
                name = "n%d_%s" % (v, _cname(c))
                if v in la.mat and v not in rootset:
                    b = la.lds[v].box
                    b = tuple(b) + ((0, 0),) * (3 - len(b))
                    UW = [self.tile[ax] + b[ax][1] - b[ax][0] for ax in range(3)]
                    cc = tuple(c) + (0,) * (3 - nd)
                    terms = ["l%d + %d" % (ax, box[ax][0] + cc[ax] - b[ax][0]) for ax in range(3)]
                    if self._m:
                        ma = self._m["ma"]
                        depth, plane, bias = self._ring(la, v)
                        parts, stride = [], 1
                        for ax in range(3):
                            if ax != ma:
                                parts.append("(%s) * %d" % (terms[ax], stride))
                                stride *= UW[ax]
                        idx = "((jz + %d) & %d) * %d + %s" % (cc[ma] + bias, depth - 1, plane,
                                                              " + ".join(parts))
                    else:
                        idx = "(%s) * %d + (%s) * %d + (%s)" % (terms[2], UW[0] * UW[1], terms[1],
                                                                UW[0], terms[0])
                    lines.append("const %s %s = s%d[%s];" % (tn, name, v, idx))
                elif n.op == "load" and self._pf is not None:
                    if any(c):
                        raise ValueError("a load phase reads field n%d at offset %s" % (v, c))
                    lines.append("const %s %s = pf%d_%d;" % (tn, name, v, self._pf))
                elif n.op == "load":
                    lines.append("const %s %s = a.ld%d[%s];" % (tn, name, self.sidx(self._ki, n.payload),
                                                                self._gaddr(c, is_out)))
                elif n.op == "imm":
                    lines.append("const %s %s = %s;" % (tn, name, c_literal(n.payload, n.ctype)))
                elif n.op == "index":
                    ax = n.payload
                    if self.open_axes[ax]:       # the cell's coordinate inside the margins
                        idx = "(g%d + %d)" % (ax, c[ax] - self.lm[ax])
                    else:
                        idx = self._wrap(ax, c[ax])
                    lines.append("const %s %s = (%s)%s;" % (tn, name, tn, idx))
                elif n.op == "size":
                    lines.append("const %s %s = (%s)%s;" % (tn, name, tn,
                                                           self._local_size(n.payload)))
                elif n.op == "bcast":
                    lines.append("const %s %s = u%d;" % (tn, name, n.args[0]))
                elif n.op == "shift":
                    c2 = tuple(x - y for x, y in zip(c, n.payload))
                    lines.append("const %s %s = n%d_%s;" % (tn, name, n.args[0], _cname(c2)))
                elif n.op == "arith":
                    op, cast_to = n.payload
                    args = ["n%d_%s" % (a_, _cname(c)) for a_ in n.args]
                    lines.append("const %s %s = %s;" % (tn, name, device_arith(
                        op, args, cast_to, [nodes[a_].ctype for a_ in n.args], n.ctype,
                        self.fast_fp32, [k.exact.get(a_) for a_ in n.args], self._comp)))
                else:
                    raise NotImplementedError(n.op)
        return lines

    def _gaddr(self, c, is_out) -> str:
        nd = self.ndim
        if is_out and all(x == 0 for x in c):
            return "addr0"
        g = [self._wrap(ax, c[ax]) for ax in range(nd)]
        g += ["0"] * (3 - nd)
        return self._addr(g)

    def _wrap(self, ax: int, off: int) -> str:
        """Coordinate g<ax> + off on the periodic grid (direct inside interior tiles)."""
        r = self._reach[ax]
        r[0], r[1] = min(r[0], off), max(r[1], off)
        self._launch_reach = max(self._launch_reach, abs(off))
        if self._direct:
            return "(g%d + %d)" % (ax, off)
        if self.open_axes[ax]:
            # only cells outside every store's region read past the grid; clamping keeps
            # those reads in bounds and their results are never stored
            return "om_clamp(g%d + %d, a.n%d)" % (ax, off, ax)
        return "om_wrap(g%d + %d, a.n%d, a.wide)" % (ax, off, ax)

    def _local_size(self, ax: int) -> str:
        m = self.lm[ax] + self.um[ax]
        return "(a.n%d - %d)" % (ax, m) if m else "(a.n%d)" % ax

    def _region_cond(self, key) -> Optional[str]:
        """C condition that (g0, g1, g2) lies in the target's region, or None if everywhere."""
        lb, ub = self.regions.get(key, ((0,) * self.ndim, (0,) * self.ndim))
        conds = []
        for ax in range(self.ndim):
            if lb[ax]:
                conds.append("g%d >= %d" % (ax, lb[ax]))
            if ub[ax]:
                conds.append("g%d < a.n%d - %d" % (ax, ax, ub[ax]))
        return " && ".join(conds) if conds else None

    # ------------------------------------------------------------ machine
    def emit(self) -> str:
        out = ["// Generated by paraiso_amd for Orthotope Machine '%s' (gfx950)." % self.name,
               "#include <math.h>", "#include <string.h>",
               "#include \"om_device.h\"",
               "#include \"paraiso_hip.h\"", ""]
        launchers = []
        kdescs = []
        for ki, sch in enumerate(self.schedules):
            ldesc = []
            self._use_form(ki)
            for la in sch.launches:
                ksrc, hsrc = self.emit_launch(ki, sch, la)
                out.append(ksrc)
                launchers.append(hsrc)
                tile = list(self.tile) if la.targets else [0, 0, 0]
                if self.batched:
                    fn = self.fn_name(ki, la.index)
                    ldesc.append("{launch_%s, {%d, %d, %d}, %d, pack_%s, sizeof(Args_%s), "
                                 "launch_batched_%s}" % (fn, tile[0], tile[1], tile[2],
                                                         len(la.reduces), fn, fn, fn))
                else:
                    ldesc.append("{launch_%s, {%d, %d, %d}, %d, nullptr, 0, nullptr}" % (
                        self.fn_name(ki, la.index), tile[0], tile[1], tile[2], len(la.reduces)))
            stored = sorted(set(s for s, _ in sch.array_stores))
            out_l = "static const pr_launch_desc k_launch_%d[] = {%s};" % (ki, ", ".join(ldesc))
            launchers.append(out_l)
            if stored:
                launchers.append("static const int k_stored_%d[] = {%s};" % (
                    ki, ", ".join(str(s) for s in stored)))
            loaded = sorted(set(n.payload for n in sch.kernel.nodes
                                if n.op == "load" and n.payload < self.n_user))
            written = sorted(set(s for s, _ in sch.kernel.stores))
            for nm, xs in (("loaded", loaded), ("written", written)):
                if xs:
                    launchers.append("static const int k_%s_%d[] = {%s};" % (
                        nm, ki, ", ".join(str(x) for x in xs)))
            kdescs.append('{"%s", %d, k_launch_%d, %d, %s, %d, %s, %d, %s}' % (
                sch.kernel.name, len(sch.launches), ki, len(stored),
                ("k_stored_%d" % ki) if stored else "nullptr",
                len(loaded), ("k_loaded_%d" % ki) if loaded else "nullptr",
                len(written), ("k_written_%d" % ki) if written else "nullptr"))
        out.extend(launchers)
        sdesc = ['{"%s", %s, %s, %d}' % (s.name, "PR_ARRAY" if s.dyn.realm is Realm.Array else "PR_SCALAR",
                                         PR_CTYPE[s.dyn.ctype], 1 if i >= self.n_user else 0)
                 for i, s in enumerate(self.statics)]
        out.append("static const pr_static_desc k_statics[] = {%s};" % ", ".join(sdesc))
        out.append("static const pr_kernel_desc k_kernels[] = {%s};" % ", ".join(kdescs))
        out.append('static const pr_machine_desc k_desc = {PR_ABI_VERSION, "%s", %d, %d, k_statics, '
                   '%d, k_kernels, %d, %d, {%d, %d, %d}, {%d, %d, %d}};'
                   % ((self.name, self.ndim, self.ns, len(self.schedules), len(self.slot_index), 1)
                      + self.lm + self.um))
        out.append('extern "C" const pr_machine_desc* pm_describe(void) { return &k_desc; }')
        return "\n".join(out) + "\n"
