"""Tile-shape autotuner for generated machines.

Replaces the reference's genetic tuner (``Language/Paraiso/Tuning/Genetic.hs``,
``examples-old/GA``), whose genome was the CUDA grid size plus Manifest and
``__syncthreads`` bits and whose fitness was cells*steps/second.  Here the
search space is the workgroup tile of the fused launches (which fixes halo
overhead, LDS footprint and hence occupancy), the fitness is measured device
time per step (HIP events on the machine's stream), and the search is
exhaustive over a small candidate set -- each candidate is a separate
compilation, so candidates are compiled ahead of time on the host and only
timed on the GPU.

    python -m paraiso_amd.autotune prepare --variant hydro_f32
    python -m paraiso_amd.autotune measure --variant hydro_f32 --size 2048 2048 --out tune.json
    python -m paraiso_amd.autotune select tune.json          # records paraiso_amd/tuned.json

Two more dimensions: ``measure --depths proceed:1 proceed_x4:4 proceed_x8:8`` times the
kernels that fuse several steps per launch (halo depth) as candidates of their own, per
step; ``alloc-prepare`` / ``alloc-measure`` / ``alloc-select`` flip the Manifest/Delayed
choice of single values nearest the scheduler's threshold and keep a flip that pays (the
per-node genome of Tuning/Genetic.hs:138-146, searched locally from the tuned machine).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
from typing import Dict, List, Tuple

HERE = os.path.dirname(os.path.abspath(__file__))
TUNED = os.path.join(HERE, "tuned.json")

# (tile, work-items per workgroup[, materialisation threshold]).  More work-items per
# workgroup raise the LDS budget, so larger tiles keep their intermediates in LDS and
# recompute less halo; the threshold is the Manifest/Delayed choice of each value
# (DecideAllocation.hs:41-43), which the reference's genetic tuner also searched.
CANDIDATES = {
    2: [((64, 8), 256), ((32, 8), 256), ((32, 16), 256), ((64, 4), 256), ((128, 4), 256),
        ((64, 16), 256), ((32, 32), 256), ((32, 16), 512), ((64, 16), 512), ((32, 32), 512),
        ((128, 8), 512), ((64, 32), 1024), ((32, 64), 1024),
        ((64, 32), 1024, 24), ((64, 32), 1024, 96), ((32, 16), 512, 24), ((32, 16), 512, 96),
        ((32, 8), 256, 16), ((32, 8), 256, 24), ((32, 16), 256, 24), ((64, 8), 256, 24),
        # streaming along y: the tile's y extent is the streaming length
        ((64, 64), 256, 48, 4), ((128, 64), 512, 48, 4), ((128, 128), 1024, 48, 8),
        ((256, 64), 1024, 48, 4), ((128, 64), 1024, 48, 8), ((64, 128), 512, 48, 8)],
    3: [((32, 4, 4), 256), ((64, 2, 2), 256), ((64, 4, 2), 256), ((64, 4, 4), 256),
        ((32, 8, 4), 256), ((64, 8, 2), 256), ((32, 8, 8), 512), ((64, 4, 4), 512),
        ((64, 8, 4), 512), ((64, 8, 8), 1024), ((32, 16, 8), 1024), ((32, 8, 16), 1024),
        ((64, 16, 4), 1024),
        # streaming along z
        ((32, 8, 64), 256, 48, 1), ((64, 8, 64), 512, 48, 1), ((64, 16, 64), 1024, 48, 1),
        ((32, 16, 64), 512, 48, 1), ((64, 8, 128), 512, 48, 2), ((32, 8, 128), 256, 48, 1),
        ((64, 8, 128), 512, 48, 4), ((64, 16, 128), 1024, 48, 2), ((128, 8, 128), 1024, 48, 2),
        ((64, 8, 256), 512, 48, 2), ((64, 4, 128), 256, 48, 2), ((128, 4, 128), 512, 48, 2),
        ((64, 16, 256), 1024, 48, 2), ((64, 8, 256), 512, 48, 4), ((32, 8, 256), 256, 48, 2),
        ((64, 16, 256), 1024, 48, 4), ((32, 8, 256), 256, 48, 4)],
}


# Candidates for machines with several kernels of different weight (an adjoint per stage):
# each kernel keeps the launch form that ran it fastest (select --per-kernel).
CANDIDATE_SETS = {
    "per_kernel_2d": [((32, 8), 256), ((16, 16), 256), ((64, 4), 256), ((64, 8), 256),
                      ((32, 16), 256), ((32, 16), 512), ((64, 8), 512), ((64, 16), 512),
                      ((32, 32), 512), ((64, 16), 1024), ((32, 32), 1024), ((64, 32), 1024),
                      ((128, 8), 1024), ((32, 8), 256, 24), ((32, 8), 256, 96),
                      ((64, 8), 256, 96)],
    # the same with the join of shared expansions (a sixth element), beside the forms the
    # staged Hydro adjoint had tuned without it
    "per_kernel_join": [((64, 8), 512, 48, 0, None, True), ((32, 16), 512, 48, 0, None, True),
                        ((32, 32), 512, 48, 0, None, True), ((16, 32), 512, 48, 0, None, True),
                        ((64, 16), 1024, 48, 0, None, True),
                        ((32, 32), 1024, 48, 0, None, True), ((64, 8), 512), ((32, 32), 512),
                        ((64, 16), 1024), ((32, 32), 1024), ((32, 16), 512, 48, 0, True, True),
                        ((32, 32), 1024, 48, 0, True, True), ((64, 8), 512, 48, 0, True, True),
                        ((32, 16), 512, 24, 0, None, True), ((32, 16), 512, 96, 0, None, True)],
    # timed with --depths: each tile at every number of steps fused per launch
    "life_depth": [((64, 32), 1024), ((128, 32), 1024), ((64, 64), 1024), ((32, 64), 1024),
                   ((128, 16), 512), ((64, 16), 512)],
    # the headline step with and without the resident tile loop (two workgroups per CU up
    # to 512 work-items and 80 KiB of LDS, one above)
    "headline_persist": [((64, 32), 1024, 48, 0, True), ((64, 32), 1024, 48, 0, False),
                         ((32, 32), 512, 48, 0, True), ((64, 16), 512, 48, 0, True),
                         ((32, 16), 512, 24, 0, True), ((32, 64), 1024, 48, 0, True),
                         ((128, 16), 1024, 48, 0, True), ((64, 32), 512, 48, 0, True)],
    "wave_depth": [((64, 16, 256), 1024, 48, 4), ((64, 8, 256), 512, 48, 4),
                   ((32, 8, 256), 256, 48, 4), ((64, 16, 256), 1024, 48, 2),
                   ((32, 16, 256), 512, 48, 4), ((64, 8, 256), 512, 48, 2),
                   ((32, 16, 256), 512, 48, 2), ((64, 16, 128), 1024, 48, 2),
                   ((128, 8, 256), 1024, 48, 2)],
}


def tuned() -> Dict:
    if os.path.exists(TUNED):
        with open(TUNED) as f:
            return json.load(f)
    return {}


def candidate_name(variant: str, tile, threads: int = 256, mat_gain: int = 48,
                   march: int = 0, persist=None, join: bool = False) -> str:
    g = "" if mat_gain == 48 else "_g%d" % mat_gain
    mr = "_m%d" % march if march else ""
    pr = "" if persist is None else ("_p" if persist else "_np")
    jn = "_j" if join else ""
    return "%s__t%s_n%d%s%s%s%s" % (variant, "x".join(str(t) for t in tile), threads, g, mr, pr,
                                    jn)


def _unpack(c):
    """(tile, threads[, mat_gain[, march[, persist[, join]]]]) -> (tile, threads, mat_gain,
    march, persist, join); persist None keeps the variant's own choice of the resident tile
    loop."""
    return (tuple(c[0]), int(c[1]), int(c[2]) if len(c) > 2 else 48,
            int(c[3]) if len(c) > 3 else 0, None if len(c) <= 4 or c[4] is None else bool(c[4]),
            bool(c[5]) if len(c) > 5 else False)


def _build_candidate(variant: str, c):
    from . import variants
    from .generator import build_machine
    mk, setup, precise = variants.VARIANTS[variant]
    tile, nt, gain, march, persist, join = _unpack(c)
    name = candidate_name(variant, tile, nt, gain, march, persist, join)
    form = dict(tile=tile, threads=nt, mat_gain=gain, march=march, join=join)
    if persist is not None:
        form["persist"] = persist
    try:
        return c, build_machine(setup.with_(**form), mk(), name, precise=precise)
    except (ValueError, subprocess.CalledProcessError) as e:
        print("candidate %s skipped: %s" % (name, str(e).splitlines()[0]), file=sys.stderr)
        return c, None


def prepare(variant: str, candidates=None, jobs: int = 4):
    """Compile the candidates in ``jobs`` worker processes; returns [(candidate, library)]."""
    from concurrent.futures import ProcessPoolExecutor
    from . import variants
    mk, setup, precise = variants.VARIANTS[variant]
    cands = candidates or CANDIDATES[setup.ndim]
    with ProcessPoolExecutor(max_workers=max(1, jobs)) as ex:
        futs = [ex.submit(_build_candidate, variant, c) for c in cands]
        return [r for r in (f.result() for f in futs) if r[1] is not None]


def _setup_inputs(variant: str, m, size):
    import numpy as np
    from . import inputs
    if variant.startswith("hydro"):
        nx, ny = size
        dt = np.float32 if "f32" in variant else np.float64
        for k, v in inputs.hydro_scalars(nx, ny).items():
            m.set(k, v)
        if "adj" in variant:
            m.set("dt", 1e-4)
        for name, arr in zip(("density", "velocity0", "velocity1", "pressure"),
                             inputs.kh_init(nx, ny, 1, dtype=dt)):
            m.set(name, arr)
    elif variant.startswith("wave"):
        dt = np.float32 if "f32" in variant else np.float64
        m.set("c", 1.0)
        m.set("dx", 1.0)
        m.set("dt", 0.5)
        m.set("fieldF", inputs.gaussian_pulse(tuple(reversed(size)), 1, dtype=dt))
        m.set("fieldG", 0.0)
    elif variant.startswith("life"):
        m.set("cell", inputs.life_init(tuple(reversed(size)), 1))
    else:
        raise KeyError(variant)


def measure(variant: str, size, steps: int = 50, warmup: int = 5, candidates=None,
            kernel: str = None, batch: int = 0, libs_of: str = None,
            kernels: List[str] = None, depths: List[Tuple[str, int]] = None) -> List[Dict]:
    """Device time per step of every compiled candidate; the step is ``proceed_pipelined``
    (after ``proceed_prime``) where the program has it, else ``kernel`` or ``proceed``.
    ``batch`` > 0 times a batch of that many grids of log-uniform sides on [64, 1024]
    (seed 2024) instead of one grid of ``size``; ``libs_of`` names the variant whose
    compiled candidates are timed (default: ``variant``).  ``kernels`` also times each of
    those OM kernels on its own (``per_kernel_ms``), for a launch form per kernel.
    ``depths`` [(kernel, k)] times each kernel that fuses k steps per launch (halo depth k)
    as another candidate, at ms / k per step: the halo depth is a tuning dimension."""
    from . import inputs, variants
    from .generator import GEN_DIR
    from .machine import Batch, Machine
    mk, setup, _ = variants.VARIANTS[variant]
    cands = candidates or CANDIDATES[setup.ndim]
    results = []
    for c in cands:
        tile, nt, gain, march, persist, join = _unpack(c)
        name = candidate_name(libs_of or variant, tile, nt, gain, march, persist, join)
        so = os.path.join(GEN_DIR, name, "lib%s.so" % name)
        if not os.path.exists(so):
            continue
        print("tune %s tile=%s threads=%d mat_gain=%d march=%d persist=%s join=%s"
              % (variant, tile, nt, gain, march, persist, join), flush=True)
        if batch:
            sizes = [tuple(int(x) for x in sz) for sz in inputs.batched_sizes(batch, 2024)]
            m = Batch(so, sizes)
            for g, sz in enumerate(sizes):
                _setup_inputs(variant, m.grids[g], sz)
        else:
            m = Machine(so, size)
            _setup_inputs(variant, m, size)
        step = kernel
        if step is None:
            step = "proceed_pipelined" if "proceed_pipelined" in m.kernels else "proceed"
            if step == "proceed_pipelined":
                m.run("proceed_prime")
        for kern, per in (depths or [(step, 1)]):
            m.run(kern, warmup)
            m.sync()
            launches = max(1, steps // per)
            ms = m.run_timed(kern, launches)
            rec = {"tile": list(tile), "threads": nt, "mat_gain": gain, "march": march,
                   "persist": persist, "join": join, "kernel": kern, "steps_per_launch": per,
                   "ms_per_step": sum(ms) / launches / per,
                   "launch_ms": [x / launches for x in ms]}
            if kernels:
                rec["per_kernel_ms"] = {}
                for k in kernels:
                    m.run(k, 2)
                    rec["per_kernel_ms"][k] = sum(m.run_timed(k, steps)) / steps
            results.append(rec)
        m.close()
    return results


# ---------------------------------------------------------------- per-node allocation choice
def alloc_candidates(variant: str, kernel: str, count: int = 8) -> List[Tuple[str, int]]:
    """Flips of the Manifest/Delayed choice nearest the scheduler's threshold for ``kernel``
    at the variant's tuned launch form: materialised values of the least gain (flip to
    Delayed: expanded in registers) and expanded values of the most (flip to Manifest: kept
    in LDS).  Value ids are those of the lowered kernel."""
    from . import variants
    from .generator.hip.ir import lower_kernel
    from .generator.hip.schedule import lds_budget, schedule_kernel
    from .generator.plan import translate
    mk, _, precise = variants.VARIANTS[variant]
    setup = variants.setup_of(variant)
    plan = translate(setup, mk())
    om = plan.om
    ki = [k.name for k in om.kernels].index(kernel)
    f = setup.form_of(kernel)
    lk = lower_kernel(om, ki, tuple(True for _ in range(om.ndim)), exact=not precise)
    sch = schedule_kernel(lk, tuple(f["tile"]), lds_budget(f["threads"]), f["mat_gain"],
                          f["march"], f.get("manifest", ()), f.get("delayed", ()),
                          join=f.get("join", False))
    mat, expanded = [], []
    for la in sch.launches:
        for gain, v, in_m in la.choices:
            (mat if in_m else expanded).append((gain, v))
    mat.sort()
    expanded.sort(reverse=True)
    flips = []
    for i in range(count):
        src = mat if i % 2 == 0 else expanded
        if not src:
            src = mat or expanded
        if not src:
            break
        gain, v = src.pop(0)
        flips.append(("delayed" if src is mat else "manifest", v))
    return flips


def _flip_name(variant: str, kernel: str, flip) -> str:
    return "%s__alloc_%s_%s%d" % (variant, kernel, flip[0][0], flip[1])


def _flipped_setup(variant: str, kernel: str, flip):
    from . import variants
    setup = variants.setup_of(variant)
    forms = dict(setup.kernel_forms)
    f = dict(forms.get(kernel, {}))
    man, dly = set(f.get("manifest", ())), set(f.get("delayed", ()))
    if flip[0] == "manifest":
        man.add(flip[1])
        dly.discard(flip[1])
    else:
        dly.add(flip[1])
        man.discard(flip[1])
    f["manifest"], f["delayed"] = tuple(sorted(man)), tuple(sorted(dly))
    forms[kernel] = f
    return setup.with_(kernel_forms=tuple(sorted(forms.items())))


def _build_flip(variant: str, kernel: str, flip):
    from . import variants
    from .generator import build_machine
    mk, _, precise = variants.VARIANTS[variant]
    name = _flip_name(variant, kernel, flip)
    try:
        return flip, build_machine(_flipped_setup(variant, kernel, flip), mk(), name,
                                   precise=precise)
    except (ValueError, subprocess.CalledProcessError) as e:
        print("flip %s skipped: %s" % (name, str(e).splitlines()[0]), file=sys.stderr)
        return flip, None


def prepare_alloc(variant: str, kernel: str, count: int = 8, jobs: int = 4):
    """Compile one machine per flip of the Manifest/Delayed choice (alloc_candidates)."""
    from concurrent.futures import ProcessPoolExecutor
    flips = alloc_candidates(variant, kernel, count)
    with ProcessPoolExecutor(max_workers=max(1, jobs)) as ex:
        futs = [ex.submit(_build_flip, variant, kernel, fl) for fl in flips]
        return [r for r in (f.result() for f in futs) if r[1] is not None]


def measure_alloc(variant: str, kernel: str, size, count: int = 8, steps: int = 50,
                  warmup: int = 5, step_kernel: str = None) -> List[Dict]:
    """Device time per step of the variant as tuned and of each compiled flip."""
    from . import variants
    from .generator import GEN_DIR
    from .machine import Machine
    libs = [(None, variants.library(variant))]
    for fl in alloc_candidates(variant, kernel, count):
        name = _flip_name(variant, kernel, fl)
        so = os.path.join(GEN_DIR, name, "lib%s.so" % name)
        if os.path.exists(so):
            libs.append((fl, so))
    out = []
    for fl, so in libs:
        print("alloc %s %s flip=%s" % (variant, kernel, fl), flush=True)
        m = Machine(so, size)
        _setup_inputs(variant, m, size)
        step = step_kernel or kernel
        if step == "proceed_pipelined":
            m.run("proceed_prime")
        m.run(step, warmup)
        m.sync()
        ms = m.run_timed(step, steps)
        out.append({"flip": list(fl) if fl else None, "kernel": step,
                    "ms_per_step": sum(ms) / steps})
        m.close()
    return out


def select_alloc(path: str) -> Dict:
    """Keep the flip that ran fastest if it beat the tuned machine by more than 0.5%
    (one step of a local search over the per-node choices; the next tuning run starts from
    the recorded choices)."""
    with open(path) as f:
        data = json.load(f)
    t = tuned()
    for variant, rec in data.items():
        base = next(r for r in rec["results"] if r["flip"] is None)["ms_per_step"]
        best = min(rec["results"], key=lambda r: r["ms_per_step"])
        entry = t.setdefault(variant, {})
        log = entry.setdefault("alloc_search", [])
        log.append({"kernel": rec["kernel"], "baseline_ms": round(base, 5),
                    "flips": {("%s:%d" % tuple(r["flip"])) if r["flip"] else "none":
                              round(r["ms_per_step"], 5) for r in rec["results"]}})
        if best["flip"] and best["ms_per_step"] < 0.995 * base:
            forms = entry.setdefault("kernel_forms", {})
            f = forms.setdefault(rec["kernel"], {})
            key = best["flip"][0]
            other = "delayed" if key == "manifest" else "manifest"
            f[key] = sorted(set(f.get(key, [])) | {best["flip"][1]})
            f[other] = [v for v in f.get(other, []) if v != best["flip"][1]]
            log[-1]["kept"] = "%s:%d" % tuple(best["flip"])
    with open(TUNED, "w") as f:
        json.dump(t, f, indent=1, sort_keys=True)
    return t


def select_per_kernel(path: str) -> Dict:
    """Record, per OM kernel, the candidate launch form that ran that kernel fastest
    (``kernel_forms`` of the variant in tuned.json); the machine is then built with one
    launch form per kernel."""
    with open(path) as f:
        data = json.load(f)
    t = tuned()
    for variant, rec in data.items():
        res = [r for r in rec["results"] if "per_kernel_ms" in r]
        forms, table = {}, {}
        for k in res[0]["per_kernel_ms"]:
            best = min(res, key=lambda r: r["per_kernel_ms"][k])
            forms[k] = {"tile": best["tile"], "threads": best["threads"],
                        "mat_gain": best.get("mat_gain", 48), "march": best.get("march", 0),
                        "ms": round(best["per_kernel_ms"][k], 5)}
            if best.get("join"):
                forms[k]["join"] = True
            if best.get("persist") is not None:
                forms[k]["persist"] = best["persist"]
            table[k] = {"%s_n%d_g%d_m%d%s%s" % ("x".join(map(str, r["tile"])), r["threads"],
                                                r.get("mat_gain", 48), r.get("march", 0),
                                                "" if r.get("persist") is None
                                                else ("_p" if r["persist"] else "_np"),
                                                "_j" if r.get("join") else ""):
                        round(r["per_kernel_ms"][k], 5) for r in res}
        t[variant] = {"kernel_forms": forms, "size": rec["size"], "steps": rec["steps"],
                      "candidates_per_kernel": table}
    with open(TUNED, "w") as f:
        json.dump(t, f, indent=1, sort_keys=True)
    return t


def select(path: str) -> Dict:
    with open(path) as f:
        data = json.load(f)
    t = tuned()
    for variant, rec in data.items():
        best = min(rec["results"], key=lambda r: r["ms_per_step"])
        keep = {k: v for k, v in t.get(variant, {}).items()
                if k in ("kernel_forms", "alloc_search", "persist")}
        t[variant] = dict(keep)
        if best.get("persist") is not None:
            t[variant]["persist"] = best["persist"]
        if best.get("join"):
            t[variant]["join"] = True
        t[variant].update({"tile": best["tile"], "threads": best["threads"],
                      "mat_gain": best.get("mat_gain", 48), "march": best.get("march", 0),
                      "kernel": best.get("kernel"), "steps_per_launch": best.get("steps_per_launch", 1),
                      "ms_per_step": best["ms_per_step"],
                      "size": rec["size"], "steps": rec["steps"],
                      "candidates": {"%s_n%d_g%d_m%d%s%s" % ("x".join(map(str, r["tile"])),
                                                             r["threads"], r.get("mat_gain", 48),
                                                             r.get("march", 0),
                                                             "" if r.get("persist") is None
                                                             else ("_p" if r["persist"] else "_np"),
                                                             ("" if r.get("steps_per_launch", 1) == 1
                                                              else "_k%d" % r["steps_per_launch"])
                                                             + ("_j" if r.get("join") else "")):
                                     round(r["ms_per_step"], 5) for r in rec["results"]}})
    with open(TUNED, "w") as f:
        json.dump(t, f, indent=1, sort_keys=True)
    return t


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    p = sub.add_parser("prepare")
    p.add_argument("--variant", required=True)
    p.add_argument("--set", default=None, help="a named candidate set (CANDIDATE_SETS)")
    q = sub.add_parser("measure")
    q.add_argument("--variant", action="append", required=True)
    q.add_argument("--size", type=int, nargs="+", action="append", required=True)
    q.add_argument("--steps", type=int, default=50)
    q.add_argument("--kernel", default=None, help="kernel to time (default: the step)")
    q.add_argument("--batch", type=int, default=0, help="time a batch of this many grids")
    q.add_argument("--libs-of", default=None, help="variant whose candidates are timed")
    q.add_argument("--kernels", nargs="*", default=None,
                   help="also time each of these kernels on its own")
    q.add_argument("--set", default=None, help="a named candidate set (CANDIDATE_SETS)")
    q.add_argument("--depths", nargs="*", default=None,
                   help="kernel:k pairs (k steps fused per launch) timed per step as more "
                        "candidates, e.g. proceed:1 proceed_x4:4 proceed_x8:8")
    q.add_argument("--out", required=True)
    pa = sub.add_parser("alloc-prepare", help="compile flips of the per-node Manifest/Delayed "
                                              "choice nearest the threshold")
    pa.add_argument("--variant", required=True)
    pa.add_argument("--kernel", required=True)
    pa.add_argument("--count", type=int, default=8)
    qa = sub.add_parser("alloc-measure")
    qa.add_argument("--variant", required=True)
    qa.add_argument("--kernel", required=True)
    qa.add_argument("--size", type=int, nargs="+", required=True)
    qa.add_argument("--count", type=int, default=8)
    qa.add_argument("--steps", type=int, default=50)
    qa.add_argument("--out", required=True)
    sa = sub.add_parser("alloc-select")
    sa.add_argument("path")
    s = sub.add_parser("select")
    s.add_argument("path")
    s.add_argument("--per-kernel", action="store_true",
                   help="record the fastest candidate of each kernel (kernel_forms)")
    a = ap.parse_args(argv)
    if a.cmd == "alloc-prepare":
        for fl, so in prepare_alloc(a.variant, a.kernel, a.count):
            print(fl, so)
        return
    if a.cmd == "alloc-measure":
        res = measure_alloc(a.variant, a.kernel, a.size, a.count, a.steps)
        with open(a.out, "w") as f:
            json.dump({a.variant: {"kernel": a.kernel, "size": a.size, "results": res}}, f,
                      indent=1)
        return
    if a.cmd == "alloc-select":
        print(json.dumps(select_alloc(a.path), indent=1))
        return
    if a.cmd == "prepare":
        for c, so in prepare(a.variant, CANDIDATE_SETS.get(a.set)):
            print(c, so)
    elif a.cmd == "measure":
        out = {}
        for variant, size in zip(a.variant, a.size):
            out[variant] = {"size": size if not a.batch else ["batch", a.batch],
                            "steps": a.steps,
                            "results": measure(variant, size, a.steps, kernel=a.kernel,
                                               batch=a.batch, libs_of=a.libs_of,
                                               kernels=a.kernels,
                                               candidates=CANDIDATE_SETS.get(a.set),
                                               depths=[(d.split(":")[0], int(d.split(":")[1]))
                                                       for d in a.depths] if a.depths else None)}
            with open(a.out, "w") as f:
                json.dump(out, f, indent=1)
    elif a.per_kernel:
        print(json.dumps(select_per_kernel(a.path), indent=1))
    else:
        print(json.dumps(select(a.path), indent=1))


if __name__ == "__main__":
    main()
