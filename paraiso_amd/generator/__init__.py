"""Code generation entry points (``Language/Paraiso/Generator.hs:29-66``).

``generate(setup, om)`` returns ``[(file name, text)]`` and ``generate_io``
writes them under ``setup.directory``, as in the reference.  The supported
language is ``Language.HIP``: a ``<Name>.hip`` with the machine's kernels and
its ``pm_describe`` descriptor, and a ``<Name>.hpp`` C++ class with the
reference's interface (constructor, one method per kernel, accessors, size
queries) that drives the machine through ``include/paraiso_hip.h``.
"""
from __future__ import annotations

import hashlib
import os
import subprocess
from typing import List, Tuple

from .. import annotation as A
from ..om.graph import OM, Reduce, Store
from .native import Language, Setup, default_setup
from .plan import Plan, translate

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC_DIR = os.path.join(PKG_DIR, "csrc")
INCLUDE_DIR = os.path.join(REPO_DIR, "include")
LIB_DIR = os.path.join(PKG_DIR, "lib")
GEN_DIR = os.path.join(PKG_DIR, "_gen")

DEFAULT_TILES = {1: (256,), 2: (64, 8), 3: (64, 4, 4)}

# the max-ILP instruction scheduler: with the same instructions (results are unchanged) it
# ran every workload here 0.6-2% faster than the default (Hydro 2048^2 FP32 and FP64, the
# batched Hydro, LinearWave 512^3 two steps per launch, Life 1024^2)
HIPCC_BASE = ["hipcc", "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
              "-mllvm", "--amdgpu-sched-strategy=max-ilp",
              "-Wno-unused-variable", "-Wno-unused-result", "-Wno-unused-value"]


def _hip_sources(setup: Setup, om: OM, precise: bool = True) -> Tuple[Plan, str, str]:
    from .hip.emit import MachineEmitter
    from .hip.host import emit_host_class
    from .hip.ir import lower_kernel
    from .hip.schedule import lds_budget, schedule_kernel
    plan = translate(setup, om)
    om_opt = plan.om
    tile = tuple(setup.tile) if setup.tile else DEFAULT_TILES[om_opt.ndim]
    cyclic = tuple(b is A.Condition.Cyclic for b in setup.boundary)
    # fast builds fold floating constants exactly and multiply by them compensated
    lowered = [lower_kernel(om_opt, ki, cyclic, exact=not precise)
               for ki in range(len(om_opt.kernels))]
    regions = _target_regions(plan, lowered) if not all(cyclic) else {}
    forms, scheds = [], []
    for ki, k in enumerate(lowered):
        f = setup.form_of(om_opt.kernels[ki].name)
        f["tile"] = tuple(f["tile"]) if f["tile"] else tile
        forms.append(f)
        scheds.append(schedule_kernel(k, f["tile"], lds_budget(f["threads"]), f["mat_gain"],
                                      f["march"], f.get("manifest", ()), f.get("delayed", ()),
                                      join=f.get("join", False)))
    src = MachineEmitter(om_opt.name, om_opt.setup.static_values, om_opt.ndim, scheds, tile,
                         fast_fp32=not precise, batched=setup.batched,
                         threads=setup.threads, open_axes=tuple(not c for c in cyclic),
                         margins=(plan.lower_margin, plan.upper_margin),
                         regions=regions, march=setup.march, prefetch=setup.prefetch,
                         persist=setup.persist, forms=forms).emit()
    hpp = emit_host_class(om_opt, setup, (plan.lower_margin, plan.upper_margin))
    return plan, src, hpp


def _target_regions(plan: Plan, lowered) -> dict:
    """Memory region [lb, n - ub) over which each store and reduce argument is computed.

    The reference computes a Manifest array only inside its subkernel's bounds, and the
    rest of its storage keeps the zeros it was allocated with (``PlanTrans.hs:75-86``);
    a Store copies the whole storage into the static (``PlanTrans.hs:243-258``) and a
    Reduce folds the whole storage (``PlanTrans.hs:719``).  Its loop addresses cells
    from the plan's lower margin whenever the bounds differ from the memory size
    (``PlanTrans.hs:440-454``), so bounds that neither cover the memory nor start at
    that margin read and write outside the arrays; such programs are rejected."""
    from .plan import _valid_to_lower, _valid_to_upper
    setup, om = plan.setup, plan.om
    lm, um = plan.lower_margin, plan.upper_margin
    out = {}
    for sub in plan.subkernels:
        lb, ub = tuple(sub.lower_boundary), tuple(sub.upper_boundary)
        if any(lb + ub) and lb != tuple(lm):
            raise NotImplementedError(
                "subkernel %s of %s computes over [%s, n - %s), which the reference's addressing "
                "(origin at the plan margin %s) does not keep inside its arrays"
                % (sub.name, plan.name, list(lb), list(ub), list(lm)))

    def region(g, value):
        valids = A.to_list(A.Valid, g.nodes[value].anot)
        if not valids:
            return None
        lb, ub = _valid_to_lower(valids[0], setup), _valid_to_upper(valids[0], setup)
        full = all(x == 0 for x in lb + ub)
        if not full and tuple(lb) != tuple(lm):
            raise NotImplementedError(
                "a subkernel of %s computes over [%s, n - %s), which the reference's addressing "
                "(origin at the plan margin %s) does not keep inside its arrays"
                % (plan.name, list(lb), list(ub), list(lm)))
        return tuple(lb), tuple(ub)

    for ki, (k, lk) in enumerate(zip(om.kernels, lowered)):
        g = k.graph
        for v, node in enumerate(g.nodes):
            if isinstance(node.inst, Store):
                r = region(g, g.pre(v)[0])
                if r is not None:
                    out[(ki, "store", node.inst.static)] = r
        for n in lk.nodes:
            if n.op != "reduce":
                continue
            for ov in n.om_nodes:
                inst_id, inst = g.inst_of_value(ov)
                if isinstance(inst, Reduce):
                    r = region(g, g.pre(inst_id)[0])
                    if r is not None:
                        out[(ki, "reduce", n.id)] = r
                    break
    return out


def generate(setup: Setup, om: OM) -> List[Tuple[str, str]]:
    if setup.language is not Language.HIP:
        raise NotImplementedError("this backend generates Language.HIP only")
    _, src, hpp = _hip_sources(setup, om)
    return [(om.name + ".hip", src), (om.name + ".hpp", hpp)]


def generate_io(setup: Setup, om: OM) -> List[Tuple[str, str]]:
    out = []
    os.makedirs(setup.directory, exist_ok=True)
    for fn, text in generate(setup, om):
        path = os.path.join(setup.directory, fn)
        with open(path, "w") as f:
            f.write(text)
        out.append((path, text))
    return out


def runtime_library() -> str:
    return os.path.join(LIB_DIR, "libparaiso_hip.so")


def build_runtime(force: bool = True) -> str:
    """Compile the hand-written runtime into ``paraiso_amd/lib/libparaiso_hip.so``."""
    os.makedirs(LIB_DIR, exist_ok=True)
    so = runtime_library()
    src = os.path.join(CSRC_DIR, "pr_runtime.cpp")
    if not force and os.path.exists(so) and os.path.getmtime(so) >= os.path.getmtime(src):
        return so
    tmp = so + ".tmp%d" % os.getpid()
    cmd = HIPCC_BASE + ["-I" + INCLUDE_DIR, "-x", "hip", src, "-o", tmp]
    subprocess.run(cmd, check=True)
    os.replace(tmp, so)
    return so


def build_machine(setup: Setup, om: OM, variant: str, precise: bool = False,
                  force: bool = False) -> str:
    """Generate and compile a machine into ``paraiso_amd/_gen/<variant>/lib<variant>.so``.

    ``precise`` compiles without FP contraction and with correctly rounded
    FP32 division and square root (bit-comparable with the CPU oracle);
    otherwise multiply-adds are contracted and FP32 division and square root
    use the hardware reciprocal and square root (about 1 ulp).
    """
    _, src, hpp = _hip_sources(setup, om, precise)
    flags = list(HIPCC_BASE) + ["-I" + CSRC_DIR, "-I" + INCLUDE_DIR]
    if precise:
        flags += ["-ffp-contract=off"]
    else:
        # FP32 division and square root use v_rcp_f32 / v_sqrt_f32 in the generated source
        flags += ["-ffp-contract=fast"]
    d = os.path.join(GEN_DIR, variant)
    os.makedirs(d, exist_ok=True)
    hip_path = os.path.join(d, variant + ".hip")
    so = os.path.join(d, "lib%s.so" % variant)
    key = hashlib.sha1((src + " ".join(flags) + open(os.path.join(CSRC_DIR, "om_device.h")).read()
                        ).encode()).hexdigest()
    stamp = os.path.join(d, "build.key")
    if (not force and os.path.exists(so) and os.path.exists(stamp)
            and open(stamp).read() == key):
        return so
    with open(hip_path, "w") as f:
        f.write(src)
    with open(os.path.join(d, om.name + ".hpp"), "w") as f:
        f.write(hpp)
    tmp = so + ".tmp%d" % os.getpid()
    subprocess.run(flags + ["-x", "hip", hip_path, "-o", tmp], check=True)
    os.replace(tmp, so)
    with open(stamp, "w") as f:
        f.write(key)
    return so


__all__ = ["Setup", "Language", "default_setup", "generate", "generate_io", "translate",
           "build_machine", "build_runtime", "runtime_library"]
