"""Launch partition, LDS materialisation and halo boxes for one OM kernel.

This is the MI355X replacement of the reference's subkernel split
(``Optimization/DependencyAnalysis.hs``, ``Generator/OMTrans.hs:136-198``) and
SASA loop (``Generator/PlanTrans.hs:406-596``).  Instead of one launch per
write group with every Manifest array round-tripping through memory, a kernel
becomes as few launches as global synchronisation allows:

* A value's *level* is the number of Reduce steps on its dependency path.
  Array values of level L are computed by launch L.  A Reduce over a level-L
  array produces a level-(L+1) scalar, finished by the last workgroup of
  launch L (the "tail"), which also evaluates the scalar arithmetic that
  depends on it and, in the last launch, commits scalar stores.
* Inside a launch every workgroup owns one tile of the grid.  Array values
  needed at several stencil offsets and expensive enough to be worth it are
  *materialised*: computed once per point over the tile grown by their halo
  and kept in LDS.  Everything else is expanded per point and per offset in
  registers, which is the reference's SASA scheme restricted to cheap values.
  Loads read at any non-zero offset become LDS halo tiles.
* Materialised values are ordered in phases separated by a workgroup barrier;
  each phase is one cooperative loop over its box.  Boxes grow backwards from
  the tile by the offsets at which later phases read them.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Set, Tuple

from ...om.dyn import CType, Realm
from .ir import LKernel, LNode

OUT = -1                      # the store phase context
_FORCE = (frozenset(), frozenset())   # per-node Manifest/Delayed overrides of the kernel scheduled
_JOIN = False                 # the kernel scheduled joins shared expansions into earlier phases
SCHED_DEBUG = False
CUT_GAIN = 2.0               # unlimited materialisation (with cuts) must save this factor of work
MAT_GAIN = 48                 # per-point operations a value must save to earn an LDS array
LDS_BUDGET = 64 * 1024        # bytes per 256-thread workgroup; larger launches are cut
LDS_MAX = 160 * 1024          # LDS of one CU (gfx950)
RED_SCRATCH = 16 * 8          # LDS a kernel adds per Reduce: one 8-byte slot per wave (<= 16)
JOIN_SHARE = 0.5              # share of a value's expansion an earlier phase must already evaluate
JOIN_GAIN = 0.1               # share of a launch's work a joined group must save


def lds_cap(reduces) -> int:
    """LDS a launch's materialised values may take: the CU's LDS less the reduction scratch
    the kernel declares beside them."""
    return LDS_MAX - RED_SCRATCH * len(reduces)


def ring_depth(planes: int) -> int:
    """Planes of a streaming ring, rounded up to a power of two (slots are masked)."""
    d = 1
    while d < planes:
        d *= 2
    return d


def lds_budget(threads: int) -> int:
    """LDS bytes a workgroup of ``threads`` may use: two resident workgroups per CU up to
    512 work-items (one can compute while the other waits at a barrier), one above."""
    if threads <= 256:
        return LDS_BUDGET
    return LDS_MAX // 2 if threads <= 512 else LDS_MAX


@dataclass
class Phase:
    index: int
    nodes: List[int]
    box: Tuple[Tuple[int, int], ...]       # per axis (lo, hi) offsets around the tile


@dataclass
class LdsBuf:
    node: int
    offset: int
    nbytes: int
    box: Tuple[Tuple[int, int], ...]


@dataclass
class ArrayLaunch:
    index: int
    stores: List[Tuple[int, int]]          # (static idx, value id) written this launch
    reduces: List[int]                      # reduce value ids finished in this tail
    targets: List[int]
    mat: Set[int]
    phases: List[Phase]
    lc: Dict[int, Dict[int, Set[Tuple[int, ...]]]]    # node -> context -> cursors
    lds: Dict[int, LdsBuf]
    lds_bytes: int
    loads_global: Set[int]                  # loads read straight from HBM
    tail_scalars: List[int]                 # scalar values computed in the tail
    commit: bool                            # this tail commits the scalar stores
    kernel_override: Optional[LKernel] = None   # node list with scratch reads (cut launches)
    # the Manifest/Delayed choice space of this launch: (gain, value, materialised) for every
    # computed value read at two or more points, gain = per-point operations LDS saves
    choices: List[Tuple[int, int, bool]] = field(default_factory=list)


@dataclass
class KernelSchedule:
    kernel: LKernel
    level: List[int]
    launches: List[ArrayLaunch]
    slot_scalars: List[int]                 # scalar values that live in device slots
    scalar_stores: List[Tuple[int, int]]
    array_stores: List[Tuple[int, int]]
    halo: int                               # largest |offset| a load is read at
    scratch: List[Tuple[int, CType]] = field(default_factory=list)   # (value id, type) per scratch array


def _add(c, d):
    return tuple(x + y for x, y in zip(c, d))


def _neg(v):
    return tuple(-x for x in v)


def schedule_kernel(k: LKernel, tile: Tuple[int, ...], budget: int = LDS_BUDGET,
                    mat_gain: int = MAT_GAIN, march: int = 0, force_manifest=(),
                    force_delayed=(), join: bool = False) -> KernelSchedule:
    """``march`` > 0: workgroups stream along the last axis, ``march`` planes per step, and
    keep each materialised value as a ring of planes (see the emitter).

    ``force_manifest`` / ``force_delayed``: values (ids of the lowered kernel) whose
    Manifest/Delayed choice is fixed rather than left to the ``mat_gain`` threshold -- the
    per-node ``AllocationChoice`` of the reference (DecideAllocation.hs:41-43), which its
    genetic tuner flipped (Tuning/Genetic.hs:138-146) and the autotuner here searches.

    ``join``: after the greedy, values that a later phase evaluates in registers and an
    earlier phase mostly evaluates already are materialised there as one group (see
    ``join_phases``), and a launch may take the whole CU's LDS: a launch whose registers
    already hold one workgroup per CU loses no occupancy by it."""
    global _FORCE, _JOIN
    _FORCE = (frozenset(force_manifest), frozenset(force_delayed))
    _JOIN = bool(join)
    if _JOIN:
        budget = LDS_MAX
    nd = k.ndim
    zero = tuple(0 for _ in range(nd))
    nodes = k.nodes
    # ---- levels
    level = [0] * len(nodes)
    for n in nodes:
        if n.op in ("load", "imm", "index", "size"):
            level[n.id] = 0
        elif n.op == "reduce":
            level[n.id] = level[n.args[0]] + 1
        else:
            level[n.id] = max([level[a] for a in n.args] + [0])
    array_stores = [(s, v) for s, v in k.stores if nodes[v].realm is Realm.Array]
    scalar_stores = [(s, v) for s, v in k.stores if nodes[v].realm is Realm.Scalar]
    reduces = [n.id for n in nodes if n.op == "reduce"]

    targets_by_level: Dict[int, List[int]] = {}
    for _, v in array_stores:
        targets_by_level.setdefault(level[v], []).append(v)
    for r in reduces:
        a = nodes[r].args[0]
        targets_by_level.setdefault(level[a], []).append(a)
    n_launch = 1 + max(list(targets_by_level) + [0])
    # scalar values evaluated in tails: level >= 1 and used by something live
    slot_scalars = [n.id for n in nodes if n.realm is Realm.Scalar and level[n.id] >= 1]

    launches = []
    scratch: List[Tuple[int, CType]] = []
    for L in range(n_launch):
        targets = sorted(set(targets_by_level.get(L, [])))
        stores_L = [(s, v) for s, v in array_stores if level[v] == L]
        red_L = [r for r in reduces if level[nodes[r].args[0]] == L]
        tail = [s for s in slot_scalars if level[s] == L + 1]
        launches.extend(_schedule_level(k, L, targets, stores_L, red_L, tail,
                                        L == n_launch - 1, mat_gain, tile, zero, scratch, budget,
                                        march=march))
    for i, la in enumerate(launches):
        la.index = i
    halo = 0
    for la in launches:
        for ph in la.phases:
            for lo, hi in ph.box:
                halo = max(halo, -lo, hi)
        for v, ctxs in la.lc.items():
            if nodes[v].op == "load":
                for curs in ctxs.values():
                    for c in curs:
                        halo = max(halo, max(abs(x) for x in c) if c else 0)
    return KernelSchedule(k, level, launches, slot_scalars, scalar_stores, array_stores, halo,
                          scratch)


SCRATCH_BASE = 1 << 20        # scratch arrays are numbered from here until the emitter places them


def _schedule_level(k, L, targets, stores, reduces, tail, commit, mat_gain, tile, zero, scratch,
                    budget, depth=0, march=0) -> List[ArrayLaunch]:
    """Schedule one level; cut it into consecutive launches while LDS is over budget.

    A cut after phase q stores every materialised value computed at or before
    q and read after it into a scratch array (its exact bits); the second
    launch reads those arrays as halo tiles, like loads of statics."""
    if budget is not None:
        budget = min(budget, lds_cap(reduces))
    la = _schedule_launch(k, L, targets, stores, reduces, tail, commit, mat_gain, tile, zero, budget,
                          march=march)
    la.kernel_override = k
    inner = [ph for ph in la.phases[:-1] if ph.index > 0]
    if la.lds_bytes > budget and (len(inner) < 2 or depth > 6):
        # cannot be cut further: keep to what the budget allows
        la = _schedule_launch(k, L, targets, stores, reduces, tail, commit, mat_gain, tile, zero,
                              budget, allow_free=False, march=march)
        la.kernel_override = k
        inner = [ph for ph in la.phases[:-1] if ph.index > 0]
    if la.lds_bytes <= budget or len(inner) < 2 or depth > 6:
        if la.lds_bytes > lds_cap(reduces):
            raise ValueError("a launch of kernel %s needs %d bytes of LDS at tile %s (at most %d)"
                             % (k.name, la.lds_bytes, tuple(tile), lds_cap(reduces)))
        return [la]
    nodes = k.nodes
    # cut where half of the computed materialised values are behind us
    n_mat = sum(len(ph.nodes) for ph in inner)
    acc, q = 0, inner[0].index
    for ph in inner[:-1]:
        acc += len(ph.nodes)
        q = ph.index
        if acc * 2 >= n_mat:
            break
    phase_of = {v: ph.index for ph in la.phases[:-1] for v in ph.nodes}
    out_phase = la.phases[-1].index
    cut = []
    for ph in inner:
        if ph.index > q:
            continue
        for v in ph.nodes:
            readers = [out_phase if ctx == OUT else ctx for ctx in la.lc.get(v, {})]
            if any(r > q for r in readers):
                cut.append(v)
    if not cut and la.lds_bytes <= lds_cap(reduces):
        return [la]
    n_scratch = len(scratch)
    try:
        if not cut:
            raise ValueError("no value crosses the cut")
        # first launch: compute the cut values and store them into scratch arrays
        first_stores = []
        for v in cut:
            sidx = SCRATCH_BASE + len(scratch)
            scratch.append((v, nodes[v].ctype))
            first_stores.append((sidx, v))
        first = _schedule_level(k, L, sorted(cut), first_stores, [], [], False, mat_gain, tile,
                                zero, scratch, budget, depth + 1, march)
        # second launch: the same targets, with the cut values read back as loads
        sidx_of = {v: s for s, v in first_stores}
        nodes2 = [LNode(n.id, "load", (), n.realm, n.ctype, sidx_of[n.id]) if n.id in sidx_of
                  else n for n in nodes]
        k2 = LKernel(k.name, nodes2, k.stores, k.ndim, k.exact)
        second = _schedule_level(k2, L, targets, stores, reduces, tail, commit, mat_gain, tile,
                                 zero, scratch, budget, depth + 1, march)
        return first + second
    except ValueError:
        # a part of the cut does not fit: keep this level in one launch within the budget
        del scratch[n_scratch:]
        la = _schedule_launch(k, L, targets, stores, reduces, tail, commit, mat_gain, tile, zero,
                              budget, allow_free=False, march=march)
        la.kernel_override = k
        if la.lds_bytes > lds_cap(reduces):
            raise ValueError("a launch of kernel %s needs %d bytes of LDS at tile %s (at most %d)"
                             % (k.name, la.lds_bytes, tuple(tile), lds_cap(reduces)))
        return [la]


def _schedule_launch(k: LKernel, L: int, targets, stores, reduces, tail, commit, mat_gain,
                     tile, zero, budget=LDS_BUDGET, allow_free=True, march=0) -> ArrayLaunch:
    nodes = k.nodes
    nd = k.ndim
    # cone of array values needed by this launch
    cone: Set[int] = set()
    stack = list(targets)
    while stack:
        x = stack.pop()
        if x in cone:
            continue
        cone.add(x)
        n = nodes[x]
        if n.op in ("arith", "shift") and n.realm is Realm.Array:
            stack.extend(a for a in n.args)
    loads = set(v for v in cone if nodes[v].op == "load" and nodes[v].realm is Realm.Array)
    target_set = set(targets)
    computed = sorted(v for v in cone if nodes[v].op in ("arith", "shift")
                      and nodes[v].realm is Realm.Array)
    force_m, force_d = _FORCE
    hinted = set(v for v in computed if nodes[v].hint_manifest and v not in target_set
                 and nodes[v].op == "arith" and v not in force_d)
    hinted |= set(v for v in computed if v in force_m and nodes[v].op == "arith"
                  and v not in target_set)
    M: Set[int] = set(loads) | hinted

    def pointwise(v, M):
        """v's expansion down to the values it reads from LDS (M) holds no Shift: every load
        it reads is read at v's own point."""
        seen, st = set(), list(nodes[v].args)
        while st:
            x = st.pop()
            if x in seen or x in M:
                continue
            seen.add(x)
            n = nodes[x]
            if n.op == "shift" and n.realm is Realm.Array:
                return False
            if n.op == "arith" and n.realm is Realm.Array:
                st.extend(n.args)
        return True

    def analyse(M):
        # phases of materialised values: loads 0, others 1 + max phase they read
        reads: Dict[int, Set[int]] = {}

        def reads_of(roots):
            seen, out, st = set(), set(), list(roots)
            first = set(roots)
            while st:
                x = st.pop()
                if x in seen:
                    continue
                seen.add(x)
                if x in M and x not in first:
                    out.add(x)
                    continue
                n = nodes[x]
                if n.op in ("arith", "shift") and n.realm is Realm.Array:
                    st.extend(n.args)
            return out
        def reads_with_cursors(root):
            out: Dict[int, Set] = {}
            seen = set()
            st = [(root, zero)]
            while st:
                x, c = st.pop()
                if (x, c) in seen:
                    continue
                seen.add((x, c))
                if x in M and x != root:
                    out.setdefault(x, set()).add(c)
                    continue
                n = nodes[x]
                if n.op == "shift":
                    st.append((n.args[0], _add(c, _neg(n.payload))))
                elif n.op == "arith" and n.realm is Realm.Array:
                    st.extend((a, c) for a in n.args if nodes[a].realm is Realm.Array)
            return out

        phase: Dict[int, int] = {}
        for v in sorted(M):
            if v in loads:
                phase[v] = 0
        for v in sorted(M):
            if v in loads:
                continue
            # a value read only at offset 0 is available in registers at the same
            # point of the same loop, so it may share the reader's phase; a value computed
            # point by point from loaded fields joins the load phase itself (phase 0 then
            # stores it in place of, or beside, the fields it reads)
            p = 0 if pointwise(v, M) else 1
            for u, curs in reads_with_cursors(v).items():
                if curs != {zero}:
                    p = max(p, phase[u] + 1)
                else:
                    p = max(p, phase[u])
            phase[v] = p
        out_roots = [t for t in targets if t not in M]
        r_out = reads_of(out_roots) | set(t for t in targets if t in M)
        out_phase = 1 + max([phase[u] for u in r_out] + [0])
        members: Dict[int, List[int]] = {}
        for v, p in phase.items():
            members.setdefault(p, []).append(v)
        # cursors per (value, phase), expanding inline values only
        lc: Dict[int, Dict[int, Set]] = {}

        def expand(P, roots):
            rootset = set(roots)
            st = [(r, zero) for r in roots]
            while st:
                x, c = st.pop()
                d = lc.setdefault(x, {}).setdefault(P, set())
                if c in d:
                    continue
                d.add(c)
                if x in M and x not in rootset:
                    continue
                n = nodes[x]
                if n.op == "shift":
                    st.append((n.args[0], _add(c, _neg(n.payload))))
                elif n.op == "arith" and n.realm is Realm.Array:
                    for a in n.args:
                        if nodes[a].realm is Realm.Array:
                            st.append((a, c))
        for p in sorted(members):
            roots_p = [v for v in members[p] if v not in loads]
            if roots_p:
                expand(p, sorted(roots_p) if p == 0 else sorted(members[p]))
        expand(OUT, out_roots)
        for t in targets:
            if t in M:
                lc.setdefault(t, {}).setdefault(OUT, set()).add(zero)
        # the OUT context reads materialised targets at 0 only; drop self entries
        for v in list(lc):
            if v in M and phase.get(v) in lc[v]:
                del lc[v][phase[v]]
        return phase, out_phase, members, lc

    def local_costs(M):
        cone_m: Dict[int, Set[int]] = {}
        lcost = {}
        for v in computed:
            n = nodes[v]
            s_: Set[int] = set()
            if n.op == "arith":
                s_.add(v)
            for a in n.args:
                if a in cone_m and a not in M:
                    s_ |= cone_m[a]
            cone_m[v] = s_
            lcost[v] = len(s_)
        return lcost

    def layout(M):
        phase, out_phase, members, lc = analyse(M)
        box: Dict[int, Tuple] = {OUT: tuple((0, 0) for _ in range(nd))}
        phases = []
        for p in sorted(members, reverse=True):
            mem = sorted(members[p])
            lo = [0] * nd
            hi = [0] * nd
            for v in mem:
                for ctx, curs in lc.get(v, {}).items():
                    cb = box[ctx]
                    for c in curs:
                        for ax in range(nd):
                            lo[ax] = min(lo[ax], cb[ax][0] + c[ax])
                            hi[ax] = max(hi[ax], cb[ax][1] + c[ax])
            pb = tuple((lo[ax], hi[ax]) for ax in range(nd))
            box[p] = pb
            phases.append(Phase(p, mem, pb))
        phases.sort(key=lambda ph: ph.index)
        phases.append(Phase(out_phase, [OUT], box[OUT]))
        last_read: Dict[int, int] = {}
        for v in M:
            rp = [out_phase if ctx == OUT else ctx for ctx in lc.get(v, {})]
            last_read[v] = max(rp) if rp else phase[v]
        lds: Dict[int, LdsBuf] = {}
        live: List[LdsBuf] = []
        top = 0
        for ph in phases[:-1]:
            if not march:
                live = [b for b in live if last_read[b.node] >= ph.index]
            # (a marching launch runs every phase at every step: its rings stay live)
            for v in ph.nodes:
                vol = 1
                for ax, (lo, hi) in enumerate(ph.box):
                    vol *= (ring_depth(hi - lo + march) if (march and ax == nd - 1)
                            else tile[ax] + hi - lo)
                nbytes = (vol * nodes[v].ctype.size + 15) // 16 * 16
                off = _first_fit(live, nbytes)
                b = LdsBuf(v, off, nbytes, ph.box)
                lds[v] = b
                live.append(b)
                top = max(top, off + nbytes)
        return phase, out_phase, members, lc, phases, lds, top

    def greedy(M, limit, rounds):
        for _round in range(rounds):
            phase, out_phase, members, lc = analyse(M)
            lcost = local_costs(M)
            ranked = []
            for v in computed:
                if v in M or nodes[v].op != "arith" or v in force_d:
                    continue
                E = sum(len(c) for c in lc.get(v, {}).values())
                gain = (E - 1) * lcost[v]
                if E >= 2 and gain >= mat_gain:
                    ranked.append((gain, v))
            if not ranked:
                break
            ranked.sort(reverse=True)
            added = False
            if limit is None:
                M = set(M) | {ranked[0][1]}
                continue
            for gain, v in ranked[:4] + [x for x in ranked[4:] if x[0] >= 2 * mat_gain][:12]:
                trial = set(M) | {v}
                if layout(trial)[6] <= limit or layout(M)[6] > limit:
                    if layout(trial)[6] <= max(limit, layout(M)[6]):
                        M = trial
                        added = True
                        break
            if not added:
                break
        return M

    def free_fields(M):
        """Materialise values computed point by point from loaded fields that take the place
        of a field in LDS (nothing else reads it): they join the load phase, so no phase,
        barrier or LDS is added and any saving pays."""
        phase, out_phase, members, lc = analyse(M)
        lcost = local_costs(M)
        top = layout(M)[6]
        for v in computed:
            if v in M or v in force_d or nodes[v].op != "arith" or not pointwise(v, M):
                continue
            E = sum(len(c) for c in lc.get(v, {}).values())
            if E < 2 or (E - 1) * lcost[v] <= 0:
                continue
            trial = set(M) | {v}
            lc_t = analyse(trial)[3]
            freed = set(u for u in M if u in loads and not any(lc_t.get(u, {}).values()))
            if SCHED_DEBUG:
                print("free_fields L%d: value %d gain %d frees %s" % (L, v, (E - 1) * lcost[v],
                                                                    sorted(freed)))
            if not freed:
                continue
            trial -= freed
            t_top = layout(trial)[6]
            if SCHED_DEBUG:
                print("free_fields L%d: value %d LDS %d -> %d" % (L, v, top, t_top))
            if t_top <= top:
                M, top = trial, t_top
                phase, out_phase, members, lc = analyse(M)
                lcost = local_costs(M)
        return M

    def work(M):
        """Per-point arithmetic a launch with materialised set M evaluates, summed over its
        phases' boxes (each evaluation of a value at a cursor counts once)."""
        _, out_phase, _, lc, phases, _, _ = layout(M)
        vol = {}
        for ph in phases:
            v_ = 1
            for ax, (lo, hi) in enumerate(ph.box):
                v_ *= tile[ax] if (march and ax == nd - 1) else tile[ax] + hi - lo
            vol[OUT if ph.nodes == [OUT] else ph.index] = v_
        total = 0
        for v, ctxs in lc.items():
            if nodes[v].op != "arith":
                continue
            for ctx, curs in ctxs.items():
                total += len(curs) * vol.get(ctx, 0)
        return total

    def join_phases(M):
        """Materialise, as one group, values that a later phase evaluates in registers and
        that an earlier phase P could compute (they read only values of phases before P),
        when at least JOIN_SHARE of their expansion is values P evaluates already.

        A reverse sweep has many such values: the cotangents of a face's left and right
        states, each read once by the cell update, expand into the face's whole forward
        computation, which the phase that materialises the face's other cotangents also
        evaluates.  One of them alone saves little, as its siblings still need the shared
        expansion; together the later phase keeps none of it.  The group is the maximal
        such values (none lies in another one's expansion).  ``work``, which counts every
        evaluation over the phases' boxes, decides: the group is kept when it falls by
        JOIN_GAIN, after leaving out the members whose absence costs least until the launch
        fits the CU's LDS."""
        phase, out_phase, members, lc = analyse(M)
        cone_sets: Dict[int, Set[int]] = {}
        reads: Dict[int, Set[int]] = {}        # materialised values each expansion reads
        for v in computed:
            n = nodes[v]
            s_ = {v} if n.op == "arith" else set()
            r_: Set[int] = set()
            for a in n.args:
                if a in M:
                    r_.add(a)
                elif a in cone_sets:
                    s_ |= cone_sets[a]
                    r_ |= reads[a]
            cone_sets[v] = s_
            reads[v] = r_
        evaluated: Dict[int, Set[int]] = {}
        for x, ctxs in lc.items():
            for ctx in ctxs:
                if ctx != OUT and ctx > 0:
                    evaluated.setdefault(ctx, set()).add(x)
        cands: Dict[int, int] = {}
        for v in computed:
            if (v in M or v in target_set or v in force_d or nodes[v].op != "arith"
                    or not lc.get(v)):
                continue
            cs = cone_sets[v]
            if len(cs) < mat_gain:
                continue
            best = None
            for P, X in evaluated.items():
                # P must be able to compute v: everything v reads comes from earlier phases
                if set(lc[v]) <= {P} or any(phase[u] >= P for u in reads[v]):
                    continue
                share = len(cs & X) / len(cs)
                if share >= JOIN_SHARE and (best is None or share > best[0]):
                    best = (share, P)
            if best is not None:
                cands[v] = best[1]
        dominated = set()
        for u, P in cands.items():
            dominated |= set(x for x in cone_sets[u] if x != u and cands.get(x) == P)
        group = set(cands) - dominated
        if not group:
            return M
        trial = set(M) | group
        w0, w1 = work(M), work(trial)
        top1 = layout(trial)[6]
        if SCHED_DEBUG:
            print("join L%d: %d values, work %d -> %d, LDS %d" % (L, len(group), w0, w1, top1))
        if top1 > lds_cap(reduces):
            # over the CU's LDS: leave out the members whose absence costs least work
            cost = sorted((work(trial - {v}) - w1, v) for v in group)
            for _, v in cost:
                trial.discard(v)
                top1 = layout(trial)[6]
                if top1 <= lds_cap(reduces):
                    break
            w1 = work(trial)
            if SCHED_DEBUG:
                print("join L%d: %d values kept, work %d, LDS %d"
                      % (L, len(trial - set(M)), w1, top1))
        if w1 <= (1 - JOIN_GAIN) * w0 and top1 <= lds_cap(reduces):
            return trial
        return M

    M = free_fields(M)
    M_fit = greedy(set(M), budget, 64)
    if budget is not None and allow_free:
        # when the LDS budget stopped the greedy, compare with materialising without limit:
        # the launch is then cut into several (values crossing a cut go through HBM),
        # which pays off when the expansion left in registers is much larger
        M_free = greedy(set(M_fit), None, 4096)
        if len(M_free) > len(M_fit):
            w_fit, w_free = work(M_fit), work(M_free)
            if SCHED_DEBUG:
                print("schedule L%d: fit %d values work %d, free %d values work %d"
                      % (L, len(M_fit), w_fit, len(M_free), w_free))
            if w_fit > CUT_GAIN * w_free:
                M_fit = M_free
    M = M_fit
    # values that ended up used once gain nothing from LDS -- except a field that the load
    # phase reads anyway for a value it computes: kept, its one later read comes from LDS
    # rather than from memory a second time
    phase, out_phase, members, lc = analyse(M)
    loaded_by_phase0 = set()
    st = [a for f in members.get(0, []) if f not in loads for a in nodes[f].args]
    while st:
        x = st.pop()
        if x in loads:
            loaded_by_phase0.add(x)
        elif x not in M and nodes[x].op == "arith" and nodes[x].realm is Realm.Array:
            st.extend(nodes[x].args)
    for v in sorted(M):
        if v in loads:
            uses = sum(len(c) for c in lc.get(v, {}).values())
            nonzero = any(c != zero for cs in lc.get(v, {}).values() for c in cs)
            if uses <= 1 and not nonzero and not (uses == 1 and v in loaded_by_phase0):
                M.discard(v)
        elif v not in force_m:
            uses = sum(len(c) for c in lc.get(v, {}).values())
            if uses <= 1:
                M.discard(v)
    if _JOIN:
        M = join_phases(M)
    phase, out_phase, members, lc, phases, lds, top = layout(M)
    if top > lds_cap(reduces):
        # the tiles the launch starts from overflow LDS (a launch behind a cut may read many
        # scratch arrays; annotated Manifest values add theirs): read the least reused loads
        # straight from memory, then keep the least reused annotated values in registers
        uses = lambda v: sum(len(c) for c in lc.get(v, {}).values())
        shed = sorted((uses(v), v) for v in M if v in loads)
        shed += sorted((uses(v), v) for v in M if v not in loads and v not in target_set)
        for _, v in shed:
            M.discard(v)
            phase, out_phase, members, lc, phases, lds, top = layout(M)
            if SCHED_DEBUG:
                print("schedule L%d: shed %s value %d, LDS %d" % (L, nodes[v].op, v, top))
            if top <= lds_cap(reduces):
                break
    loads_global = set(v for v in loads if v not in M)
    # the choice space the autotuner searches: computed values read at two or more points
    lcost = local_costs(M)
    choices = []
    for v in computed:
        if nodes[v].op != "arith" or v in target_set:
            continue
        E = sum(len(c) for c in lc.get(v, {}).values())
        if E >= 2 or v in M:
            choices.append(((max(E, 2) - 1) * lcost[v], v, v in M))
    return ArrayLaunch(L, stores, reduces, targets, M, phases, lc, lds, top, loads_global,
                       tail, commit, choices=sorted(choices, reverse=True))


def _first_fit(live: List[LdsBuf], nbytes: int) -> int:
    spans = sorted((b.offset, b.offset + b.nbytes) for b in live)
    pos = 0
    for a, e in spans:
        if pos + nbytes <= a:
            return pos
        pos = max(pos, e)
    return pos
