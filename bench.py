#!/usr/bin/env python3
"""Throughput of the explicit-stencil hot path on one MI355X.

Workload (the headline): 2-D compressible Euler (examples/Hydro/HydroMain.hs),
gamma = 1.4, 2048 x 2048 periodic, Kelvin-Helmholtz initial state, CFL 0.4,
FP32 bulk arithmetic with the CFL minimum reduced in FP64.  A step is one call
of the OM kernel ``proceed`` (both HLLC stages plus the CFL reduction).

Printed as one JSON line: Gcell-updates/s, the roofline of the dominant
kernel (HIP events on the machine's own stream), the same run in FP64, the
FP32-vs-FP64 error, and the CPU path (the oracle's restatement of the
reference's generated C++, OpenMP on the host cores) on a bounded sample.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# the CPU baseline's OpenMP threads stay on neighbouring cores; read when libgomp loads
os.environ.setdefault("OMP_PROC_BIND", "close")
# CPUs this process may use, before an OpenMP runtime binds the main thread to one of them
AFFINITY_CPUS = len(os.sched_getaffinity(0))

HBM_PEAK_GBPS = 8000.0
FIELDS = ("density", "velocity0", "velocity1", "pressure")


def setup_hydro(m, nx, ny, seed, dtype):
    from paraiso_amd import inputs
    for k, v in inputs.hydro_scalars(nx, ny, cfl=0.4).items():
        m.set(k, v)
    for name, arr in zip(FIELDS, inputs.kh_init(nx, ny, seed, dtype=dtype)):
        m.set(name, arr)


def conserved(m, gamma=1.4):
    import numpy as np
    r, u, v, p = (m.get(f).astype(np.float64) for f in FIELDS)
    return np.stack([r, r * u, r * v, p / (gamma - 1) + 0.5 * r * (u * u + v * v)])


def time_machine(variant, nx, ny, steps, warmup, seed, dtype, bytes_per_cell,
                 kernel="proceed_pipelined", prime="proceed_prime"):
    """Warm up, then time `steps` calls of `kernel` replayed as one HIP graph (wall clock
    between device synchronisations); another `steps` calls with an event pair around every
    launch give the device time of each launch.

    The default step is ``proceed_pipelined``: ``proceed`` with the CFL minimum for the
    next step reduced in this step's store phase (``prime`` computes the first one
    before the warmup); it stores bit for bit what ``proceed`` stores."""
    import torch
    from paraiso_amd import variants
    m = variants.machine(variant, (nx, ny))
    setup_hydro(m, nx, ny, seed, dtype)
    if prime:
        m.run(prime)
    m.run(kernel, warmup)
    m.run(kernel, steps, graph=True)             # builds and warms the graph
    m.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    m.run(kernel, steps, graph=True)             # the timed steps: one HIP graph
    m.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    launch_ms = m.run_timed(kernel, steps)       # device time per launch (events)
    cells = nx * ny
    dom = max(range(len(launch_ms)), key=lambda l: launch_ms[l])
    avg_ms = launch_ms[dom] / steps
    achieved = cells * bytes_per_cell / (avg_ms * 1e-3) / 1e9
    return m, {
        "value": cells * steps / wall / 1e9,
        "ms_per_step": wall / steps * 1e3,
        "launch_ms_per_step": [x / steps for x in launch_ms],
        "dominant_launch": dom,
        "dominant_avg_ms": avg_ms,
        "achieved_GBps": achieved,
    }


VALU_PEAK = 256 * 4 * 2.4e9 / 4      # wave64 VALU instructions/s, gfx950 without packed FP32
LANE_VALU_PEAK = VALU_PEAK * 64       # lane instructions/s: 256 CUs x 4 SIMDs x 16 lanes
PROFILE = "profiles/hydro2d_kh_2048x2048_fp32.md"


def useful_ops_per_cell(variant, kernel):
    """Operations per cell of an OM kernel as written, each value once per cell (the lowered
    IR, after value numbering and constant folding): floating arithmetic (add, sub, mul,
    div, sqrt, min, max, abs, neg) and compares, selects and logic.  This is the work a cell
    update needs without halo recomputation, address arithmetic or data movement."""
    from paraiso_amd import variants
    from paraiso_amd.generator.hip.ir import lower_kernel
    from paraiso_amd.generator.plan import translate
    from paraiso_amd.om.arith import COMPARISONS, Op
    from paraiso_amd.om.dyn import Realm
    mk, _, precise = variants.VARIANTS[variant]
    plan = translate(variants.setup_of(variant), mk())
    om = plan.om
    ki = [k.name for k in om.kernels].index(kernel)
    lk = lower_kernel(om, ki, (True,) * om.ndim, exact=not precise)
    fp = {Op.Add, Op.Sub, Op.Mul, Op.Div, Op.Inv, Op.Neg, Op.Sqrt, Op.Abs, Op.Max, Op.Min}
    logic = set(COMPARISONS) | {Op.Select, Op.And, Op.Or, Op.Not}
    out = {"fp_arith": 0, "compare_select": 0}
    for n in lk.nodes:
        if n.op != "arith" or n.realm is not Realm.Array:
            continue
        op = n.payload[0]
        if op in fp:
            out["fp_arith"] += 1
        elif op in logic:
            out["compare_select"] += 1
    return out


def build_key(variant):
    """Hash of the generated source and compiler flags of a built machine."""
    from paraiso_amd.generator import GEN_DIR
    path = os.path.join(GEN_DIR, variant, "build.key")
    return open(path).read().strip() if os.path.exists(path) else None


def profile_figures(variant, kernel):
    """HBM bytes and VALU instructions per launch of `kernel`, from the rocprofv3 summary
    committed under profiles/ for this workload, if it was taken with the same tile."""
    path = os.path.join(ROOT, PROFILE)
    if not os.path.exists(path):
        return None, None
    from paraiso_amd import variants
    rec = {}
    with open(path) as f:
        for line in f:
            for key in ("profile_config", "traffic_bytes_per_launch",
                        "valu_wave_instructions_per_launch", "valu_active_quad_cycles_per_launch",
                        "valu_transcendental_per_launch", "valu_mix_per_launch"):
                if line.startswith(key + ":"):
                    rec[key] = json.loads(line.split(":", 1)[1])
    cfg = rec.get("profile_config", {})
    setup = variants.setup_of(variant)
    if (cfg.get("variant") != variant or list(setup.tile or ()) != cfg.get("tile")
            or setup.threads != cfg.get("threads")
            or bool(setup.persist) != bool(cfg.get("persist", False))
            or cfg.get("build_key") != build_key(variant)):
        return None, None
    valu = {k: rec.get("valu_%s_per_launch" % k, {}).get(kernel)
            for k in ("wave_instructions", "active_quad_cycles", "transcendental", "mix")}
    return (rec.get("traffic_bytes_per_launch", {}).get(kernel),
            valu if valu["wave_instructions"] is not None else None)


def accuracy(nx, ny, steps=500, seed=1):
    """FP32 against FP64 after `steps` steps from the same KH state: relative L1 of the
    conserved state and per field (L1 and Linf), and drift of total mass and energy.

    Two FP32 machines are compared with the FP64 one: the product machine (``hydro_f32``:
    FP32 bulk arithmetic, constants folded exactly, compensated products by the constants
    whose reciprocal the kernel also multiplies by) and the reference's own FP32 program
    correctly rounded (``hydro_f32_precise``, which the tests hold bit for bit equal to the
    FP32 oracle over 500 steps at 512^2)."""
    import numpy as np
    from paraiso_amd import variants
    out, totals = {}, {}
    ms = {}
    for name, dt in (("hydro_f32", np.float32), ("hydro_f32_precise", np.float32),
                     ("hydro_f64", np.float64)):
        m = variants.machine(name, (nx, ny))
        setup_hydro(m, nx, ny, seed, dt)
        c0 = conserved(m)
        m.run("proceed_prime")
        m.run("proceed_pipelined", steps)
        ms[name] = m
        c = conserved(m)
        totals[name] = {"mass_drift": float(abs(c[0].sum() - c0[0].sum()) / c0[0].sum()),
                        "energy_drift": float(abs(c[3].sum() - c0[3].sum()) / c0[3].sum())}
    a, r, b = ms["hydro_f32"], ms["hydro_f32_precise"], ms["hydro_f64"]
    ca, cr, cb = conserved(a), conserved(r), conserved(b)
    out["steps"] = steps
    out["rel_l1_conserved"] = float(np.abs(ca - cb).sum() / np.abs(cb).sum())
    out["reference_fp32_program_rel_l1_conserved"] = float(np.abs(cr - cb).sum()
                                                           / np.abs(cb).sum())
    for f in FIELDS:
        x, y = a.get(f).astype(np.float64), b.get(f)
        out[f] = {"rel_l1": float(np.abs(x - y).sum() / np.abs(y).sum()),
                  "linf": float(np.abs(x - y).max())}
    out["time_f32"], out["time_f64"] = float(a.get("time")), float(b.get("time"))
    out["drift_f32"], out["drift_f64"] = totals["hydro_f32"], totals["hydro_f64"]
    out["drift_reference_fp32_program"] = totals["hydro_f32_precise"]
    for m in ms.values():
        m.close()
    return out


def hbm_copy_gbps():
    import torch
    n = 1 << 28                                # 1 GiB of float32 each way
    a = torch.empty(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(10):
        b.copy_(a)
    e.record()
    torch.cuda.synchronize()
    return 2 * n * 4 * 10 / (s.elapsed_time(e) * 1e-3) / 1e9


def bench_batched(n_grids=256, steps=200, seed=2024):
    """256 Hydro grids, sides log-uniform on [64, 1024], FP32, one HIP graph of `steps` steps
    of the carried-CFL step (each a single launch over the tiles of every grid)."""
    import numpy as np
    import torch
    from paraiso_amd import inputs, variants
    sizes = [tuple(int(x) for x in s) for s in inputs.batched_sizes(n_grids, seed)]
    b = variants.batch("hydro_f32_batch", sizes)
    for g, (nx, ny) in enumerate(sizes):
        setup_hydro(b.grids[g], nx, ny, seed + g, np.float32)
    b.run("proceed_prime")
    b.run("proceed_pipelined", 2, graph=False)
    b.run("proceed_pipelined", steps, graph=True)   # builds and warms the graph
    b.sync()
    for g, (nx, ny) in enumerate(sizes):
        setup_hydro(b.grids[g], nx, ny, seed + g, np.float32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.run("proceed_prime")                          # the first step's CFL minimum
    b.run("proceed_pipelined", steps, graph=True)
    b.sync()
    t1 = time.perf_counter()
    cells = b.cells()
    # a seeded sample of grids against their own single-grid runs of the same steps
    pick = sorted(int(i) for i in np.random.default_rng(seed).choice(n_grids, 4, replace=False))
    worst, same = 0.0, True
    for g in pick:
        nx, ny = sizes[g]
        m = variants.machine("hydro_f32_batch", (nx, ny))
        setup_hydro(m, nx, ny, seed + g, np.float32)
        m.run("proceed_prime")
        m.run("proceed_pipelined", steps)
        for f in FIELDS:
            a, c = b.grids[g].get(f), m.get(f)
            same = same and bool(np.array_equal(a, c))
            worst = max(worst, float(np.abs(a.astype(np.float64) - c).max()))
        same = same and b.grids[g].get("time") == m.get("time")
        m.close()
    ms = b.run_timed("proceed_pipelined", 10)
    b.close()
    return {"workload": "hydro2d-kh-batched-%d-grids-fp32" % n_grids, "grids": n_grids,
            "cells_per_step": cells, "steps": steps,
            "value": cells * steps / (t1 - t0) / 1e9, "unit": "Gcell-updates/s",
            "ms_per_step": (t1 - t0) / steps * 1e3, "launch_ms_per_step": [x / 10 for x in ms],
            "sizes_seed": seed,
            "single_grid_check": {"grids": pick, "sizes": [list(sizes[g]) for g in pick],
                                  "bit_identical": same, "max_abs_diff": worst}}


STAGES = ["proceed_stage1", "proceed_stage2"]
STAGE_ADJOINTS = ["proceed_stage2_adjoint", "proceed_stage1_adjoint"]


def bench_adjoint(n=512, steps=200, every=20, dt=1e-4, variant="hydro_adj2_f32_fast"):
    """Fixed-dt Hydro 512^2 FP32: 200 forward steps against forward + checkpointed adjoint.

    The forward with checkpoints runs the fused step ``proceed``; the backward replays each
    segment in the step's two stages and runs one adjoint kernel per stage (stage 2, then
    stage 1), each kernel at the launch form tuned for it.  In correctly rounded builds the
    two stages store what ``proceed`` stores bit for bit; here (FP32 fast) they agree to
    rounding."""
    import numpy as np
    import torch
    from paraiso_amd import inputs, variants
    m = variants.machine(variant, (n, n))
    names = [s[0] for s in m.statics]
    snap = [x for x in names if not x.startswith("adj_") and not x.startswith("__")]
    S = (steps + every - 1) // every
    m.snapshot_alloc(S + every, snap)

    def fresh():
        setup_hydro(m, n, n, 7, np.float32)
        m.set("dt", dt)
    fresh()
    for k in ["proceed"] + STAGES + STAGE_ADJOINTS:       # first launches outside the timing
        m.run(k, 2)
    m.snapshot_save(0)                                     # the runtime's copy kernel too
    m.snapshot_load(0)
    m.sync()
    fresh()
    t0 = time.perf_counter()
    m.run("proceed", steps)
    m.sync()
    t_fwd = time.perf_counter() - t0
    fresh()
    t0 = time.perf_counter()
    m.adjoint_forward("proceed", steps, every)
    m.sync()
    t_af = time.perf_counter() - t0
    # seed: dJ/du at the final state for J = sum over [0.25,0.75]^2 of rho |v|^2 / 2
    rho, v0, v1 = (torch.tensor(m.get(f), device="cuda", dtype=torch.float64)
                   for f in ("density", "velocity0", "velocity1"))
    c = (torch.arange(n, device="cuda", dtype=torch.float64) + 0.5) / n
    inside = ((c >= 0.25) & (c <= 0.75)).to(torch.float64)
    mask = inside[:, None] * inside[None, :]
    seeds = {"adj_density": 0.5 * (v0 * v0 + v1 * v1) * mask, "adj_velocity0": rho * v0 * mask,
             "adj_velocity1": rho * v1 * mask, "adj_pressure": torch.zeros_like(rho)}
    for k, v in seeds.items():
        m.set(k, v.float().cpu().numpy())
    m.sync()
    t0 = time.perf_counter()
    m.adjoint_backward(STAGES, STAGE_ADJOINTS, steps, every, tape_after=1)
    m.sync()
    t_ab = time.perf_counter() - t0
    g = m.get("adj_density")
    kernel_us = {k: 1e3 * sum(m.run_timed(k, 20)) / 20 for k in ["proceed"] + STAGES + STAGE_ADJOINTS}
    m.close()
    # the same fixed-dt forward on the plain FP64 single-grid path (correctly rounded FP64) at
    # its fastest launch form
    m64 = variants.machine("hydro_fixed_f64", (n, n))
    setup_hydro(m64, n, n, 7, np.float64)
    m64.set("dt", dt)
    m64.run("proceed", 5)
    m64.sync()
    setup_hydro(m64, n, n, 7, np.float64)
    m64.set("dt", dt)
    t0 = time.perf_counter()
    m64.run("proceed", steps)
    m64.sync()
    t_fwd64 = time.perf_counter() - t0
    m64.close()
    return {"workload": "hydro2d-adjoint-%dx%d-fp32-fixed-dt" % (n, n), "variant": variant,
            "steps": steps, "checkpoint_every": every,
            "forward_kernel": "proceed", "replay_kernels": STAGES,
            "adjoint_kernels": STAGE_ADJOINTS,
            "fwd_ms": t_fwd * 1e3, "fwd_with_checkpoints_ms": t_af * 1e3, "adjoint_ms": t_ab * 1e3,
            "fwd_adj_over_fwd": (t_af + t_ab) / t_fwd,
            "fp64_fwd_ms": t_fwd64 * 1e3,
            "fwd_adj_over_fp64_fwd": (t_af + t_ab) / t_fwd64,
            "kernel_us_per_launch": kernel_us,
            "grad_density_l1": float(np.abs(g).sum())}


def bench_life(n=1024, gens=1000, seed=1):
    """Conway Life 1024^2 int32 periodic, Bernoulli(0.5) cells, 1000 generations: one graph of
    ``proceed_x4`` or ``proceed_x8`` launches (four or eight generations per launch, bit-identical
    to as many ``proceed``); both are timed and the faster is reported."""
    runs = [_bench_life(n, gens, seed, k) for k in (4, 8)]
    best = dict(min(runs, key=lambda r: r["ms_per_step"]))
    best["alternatives"] = {"proceed_x%d" % r["generations_per_launch"]: r["value"] for r in runs}
    return best


def _bench_life(n, gens, seed, per_launch):
    import torch
    from paraiso_amd import inputs, variants
    m = variants.machine("life_i32", (n, n))
    m.run("init")
    m.set("cell", inputs.life_init((n, n), seed))
    kern = "proceed_x%d" % per_launch
    m.run(kern, gens // per_launch, graph=True)      # builds and warms the graph
    m.run("init")
    m.set("cell", inputs.life_init((n, n), seed))
    m.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    m.run(kern, gens // per_launch, graph=True)
    m.sync()
    t1 = time.perf_counter()
    pop, gen = int(m.get("population")), int(m.get("generation"))
    ms = m.run_timed(kern, 25)
    cells = n * n
    launch_s = sum(ms) / 25 * 1e-3
    return {"workload": "life-%dx%d-int32-%d-generations" % (n, n, gens), "generations": gens,
            "value": cells * gens / (t1 - t0) / 1e9, "unit": "Gcell-updates/s",
            "ms_per_step": (t1 - t0) / gens * 1e3,
            "launch_ms_per_launch": [x / 25 for x in ms], "generations_per_launch": per_launch,
            "algorithmic_bytes_per_cell": 8,
            # a launch of k generations reads and writes each cell once: 8 B per cell moved
            "hbm_bytes_per_launch": cells * 8,
            "achieved_GBps": cells * 8 / launch_s / 1e9,
            "population": pop, "generation": gen}


def bench_wave(n=512, steps=300, every=20):
    """LinearWave 3-D, 512^3 FP32, c = 1, dt = 0.5 dx, Gaussian pulse (8 cells, seed 3).

    Then forward + checkpointed adjoint of the misfit J = 1/2 sum (f_N - d)^2 with
    d the same pulse from seed 4; the adjoint is seeded with f_N - d."""
    import numpy as np
    import torch
    from paraiso_amd import inputs, variants
    m = variants.machine("wave3d_f32", (n, n, n))
    m.set("c", 1.0)
    m.set("dx", 1.0)
    m.set("dt", 0.5)
    m.set("fieldF", inputs.gaussian_pulse((n, n, n), 3, dtype=np.float32))
    m.set("fieldG", 0.0)
    m.run("proceed", 3)
    m.sync()
    e0 = m.get("energy")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ms = m.run_timed("proceed", steps)
    m.sync()
    t1 = time.perf_counter()
    cells = n ** 3
    dom = max(range(len(ms)), key=lambda l: ms[l])
    achieved = cells * 16 / (ms[dom] / steps * 1e-3) / 1e9
    drift = abs(m.get("energy") - e0) / abs(e0)
    runs = {"proceed": {"kernel": "proceed", "value": cells * steps / (t1 - t0) / 1e9,
                        "ms_per_step": (t1 - t0) / steps * 1e3, "achieved_GBps": achieved,
                        "frac": achieved / HBM_PEAK_GBPS}}
    # k steps per launch (proceed_x<k>: k proceed steps fused, equal to rounding in FP32)
    for k in (2,):
        kern = "proceed_x%d" % k
        m.set("fieldF", inputs.gaussian_pulse((n, n, n), 3, dtype=np.float32))
        m.set("fieldG", 0.0)
        m.run(kern, 2)
        m.sync()
        t2 = time.perf_counter()
        msk = m.run_timed(kern, steps // k)
        m.sync()
        t3 = time.perf_counter()
        done = (steps // k) * k
        achk = cells * 16 / (msk[0] / (steps // k) * 1e-3) / 1e9
        runs[kern] = {"kernel": kern, "value": cells * done / (t3 - t2) / 1e9,
                      "ms_per_step": (t3 - t2) / done * 1e3, "achieved_GBps": achk,
                      "frac": achk / HBM_PEAK_GBPS}
    best = max(runs.values(), key=lambda r: r["value"])
    out = {"workload": "wave3d-%d^3-fp32" % n, "steps": steps, "unit": "Gcell-updates/s",
           "algorithmic_bytes_per_cell_per_launch": 16, "energy_drift": drift}
    out.update(best)
    out["alternatives"] = runs
    # forward + adjoint of the misfit, with the faster forward kernel and its adjoint (the
    # two-step pair runs steps / 2 launches, a checkpoint every every / 2)
    k = 2 if best["kernel"] == "proceed_x2" else 1
    fwd, adj = ("proceed_x2", "proceed_adjoint_x2") if k == 2 else ("proceed", "proceed_adjoint")
    pulse = inputs.gaussian_pulse((n, n, n), 3, dtype=np.float32)
    target = inputs.gaussian_pulse((n, n, n), 4, dtype=np.float32)
    S = (steps // k + every // k - 1) // (every // k)
    names = [x[0] for x in m.statics]
    m.snapshot_alloc(S + every // k, [x for x in names if not x.startswith(("adj_", "__"))])

    def reset():
        m.set("fieldF", pulse)
        m.set("fieldG", 0.0)
        m.sync()
    reset()
    t0 = time.perf_counter()
    m.run(fwd, steps // k)
    m.sync()
    t_fwd = time.perf_counter() - t0
    # once untimed, so that every snapshot slot has been written before (the first write to
    # fresh device memory is slower than the steady state an optimisation loop sees)
    reset()
    m.adjoint_forward(fwd, steps // k, every // k)
    reset()
    t0 = time.perf_counter()
    m.adjoint_forward(fwd, steps // k, every // k)
    m.sync()
    t_af = time.perf_counter() - t0
    r = m.get("fieldF").astype(np.float64) - target
    misfit = 0.5 * float((r * r).sum())
    m.set("adj_fieldF", r.astype(np.float32))
    m.set("adj_fieldG", 0.0)
    m.sync()
    t0 = time.perf_counter()
    m.adjoint_backward(fwd, adj, steps // k, every // k)
    m.sync()
    t_ab = time.perf_counter() - t0
    out["adjoint"] = {"misfit": misfit, "kernels": [fwd, adj], "checkpoint_every": every,
                      "fwd_ms": t_fwd * 1e3, "fwd_with_checkpoints_ms": t_af * 1e3,
                      "adjoint_ms": t_ab * 1e3, "fwd_adj_over_fwd": (t_af + t_ab) / t_fwd,
                      "grad_l2": float(np.sqrt((m.get("adj_fieldF").astype(np.float64) ** 2).sum()))}
    return out


def _cpu_info():
    """Host CPU as the CPU baseline ran on it: model name, logical CPUs, the CPUs this process
    may use, and the OpenMP threads and binding of the baseline."""
    model = "unknown"
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    affinity = AFFINITY_CPUS
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": affinity,
            "omp_threads": int(os.environ.get("OMP_NUM_THREADS", affinity)),
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND", "")}


def _median_rate(run, work, budget_s, max_steps, runs=3):
    """Median over `runs` of work done per second, each run stepping until `budget_s` or
    `max_steps`; `run(k)` does k steps, `work` is the work per step."""
    import numpy as np
    rates, steps_total, secs = [], 0, 0.0
    for _ in range(runs):
        run(1)                                   # first touch outside the timed region
        steps, t0 = 0, time.perf_counter()
        while steps < max_steps and time.perf_counter() - t0 < budget_s:
            run(1)
            steps += 1
        dt = time.perf_counter() - t0
        rates.append(work * steps / dt)
        steps_total += steps
        secs += dt
    return float(np.median(rates)), [round(x, 6) for x in rates], steps_total, secs


def cpu_baseline(nx, ny, budget_s=4.0):
    """The reference's CPU path on the host cores: the oracle's restatement of the reference's
    generated CPlusPlus class (std::vector storage, SASA loops under OpenMP, serial reduces,
    deep-copy stores), built g++ -O3 -fopenmp -ffp-contract=off, OMP_PROC_BIND=close, every
    OpenMP thread the process has.  Bounded samples (median of three runs of about
    `budget_s` each) of every workload; the headline's figure is the top-level value."""
    import numpy as np
    from fractions import Fraction
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import oracle_machine
    from oracle.prebuild import BASELINE_OPT, batched_sample
    from paraiso_amd import inputs
    from paraiso_amd.om.dyn import CType
    from paraiso_amd.schemes.hydro import hydro_om
    from paraiso_amd.schemes.life import life_om
    from paraiso_amd.schemes.linear_wave import linear_wave_om
    G = Fraction(7, 5)
    info = _cpu_info()
    flags = "g++ %s -fopenmp -ffp-contract=off" % BASELINE_OPT
    # Hydro KH 2048^2 FP32 (the headline)
    m = oracle_machine(lambda: hydro_om(CType.Float, G), (nx, ny), opt=BASELINE_OPT)
    setup_hydro(m, nx, ny, 1, np.float32)
    hyd, runs, st, secs = _median_rate(lambda k: m.run("proceed", k), nx * ny / 1e9, budget_s, 40)
    m.close()
    out = {"value": hyd, "unit": "Gcell-updates/s", "cores": info["omp_threads"], "kind": "port",
           "runs": runs, "sample": "median of 3 runs, %d steps of Hydro %dx%d FP32 in %.1f s "
           "(%s)" % (st, nx, ny, secs, flags)}
    out.update(info)
    others = {}
    # Life 1024^2
    n = 1024
    m = oracle_machine(life_om, (n, n), opt=BASELINE_OPT)
    m.run("init")
    m.set("cell", inputs.life_init((n, n), 1))
    v, runs, st, secs = _median_rate(lambda k: m.run("proceed", k), n * n / 1e9, budget_s, 1000)
    others["life"] = {"value": v, "unit": "Gcell-updates/s", "runs": runs,
                      "sample": "%d generations of Life %dx%d in %.1f s" % (st, n, n, secs)}
    m.close()
    # LinearWave 512^3 FP32 (a stated subset of the 300 steps)
    n3 = (512, 512, 512)
    m = oracle_machine(lambda: linear_wave_om(3, CType.Float), n3, opt=BASELINE_OPT)
    m.set("c", 1.0)
    m.set("dx", 1.0)
    m.set("dt", 0.5)
    m.set("fieldF", inputs.gaussian_pulse(n3, 3, dtype=np.float32))
    m.set("fieldG", 0.0)
    v, runs, st, secs = _median_rate(lambda k: m.run("proceed", k), 512 ** 3 / 1e9, budget_s, 20)
    others["wave3d"] = {"value": v, "unit": "Gcell-updates/s", "runs": runs,
                        "sample": "%d steps of LinearWave 512^3 FP32 in %.1f s" % (st, secs)}
    m.close()
    # batched Hydro: a seeded sample of the 256 grid sizes, stepped in turn
    sample = batched_sample()
    ms = []
    for g, (bx, by) in enumerate(sample):
        mg = oracle_machine(lambda: hydro_om(CType.Float, G), (bx, by), opt=BASELINE_OPT)
        setup_hydro(mg, bx, by, 2024 + g, np.float32)
        ms.append(mg)
    cells = sum(bx * by for bx, by in sample)

    def batch_step(k):
        for mg in ms:
            mg.run("proceed", k)
    v, runs, st, secs = _median_rate(batch_step, cells / 1e9, budget_s, 200)
    others["batched"] = {"value": v, "unit": "Gcell-updates/s", "runs": runs,
                         "sample": "%d steps of %d of the 256 grids (%s, %d cells) in %.1f s"
                         % (st, len(sample), sample, cells, secs)}
    for mg in ms:
        mg.close()
    # the adjoint: the CPU FP64 reference adjoint (torch autograd through the independent
    # restatement oracle/hydro_torch.py), forward + backward per step at 512^2
    others["adjoint"] = _cpu_adjoint(512, budget_s)
    out["others"] = others
    return out


def _cpu_adjoint(n, budget_s, dt=1e-4, steps=4):
    import numpy as np
    import torch
    from oracle import hydro_torch as HT
    from paraiso_amd import inputs
    torch.set_num_threads(int(os.environ.get("OMP_NUM_THREADS", AFFINITY_CPUS)))
    state = inputs.kh_init(n, n, 7)
    times = []
    t_end = time.perf_counter() + 3 * budget_s
    while len(times) < 3 and (not times or time.perf_counter() < t_end):
        t = [torch.tensor(a, dtype=torch.float64, requires_grad=True) for a in state]
        t0 = time.perf_counter()
        out, _ = HT.run(t, steps, (1.0 / n, 1.0 / n), 1.4, dt=dt)
        J, _ = HT.kinetic_energy_in_square(*out[:3])
        J.backward()
        times.append(time.perf_counter() - t0)
    per = float(np.median(times)) / steps
    return {"value": n * n / per / 1e9, "unit": "Gcell-updates/s (forward + adjoint)",
            "ms_per_step_fwd_adj": per * 1e3, "runs": [round(x, 3) for x in times],
            "sample": "%d steps of %d^2 FP64, forward and reverse through torch autograd "
                      "(oracle/hydro_torch.py), median of %d runs" % (steps, n, len(times))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=2048)
    ap.add_argument("--ny", type=int, default=2048)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batched, adjoint, 3-D wave and Life measurements")
    ap.add_argument("--extras", default="batched,adjoint,wave3d,life",
                    help="comma-separated subset of the extra measurements to run")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (grids would shard "
              "by slabs along their slowest axis with one halo exchange per launch)",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(1)
    nx, ny, K, W = args.nx, args.ny, args.steps, args.warmup
    m32, r32 = time_machine("hydro_f32", nx, ny, K, W, 1, np.float32, 32)
    _, r32u = time_machine("hydro_f32", nx, ny, K, W, 1, np.float32, 32, "proceed", None)
    out = {
        "metric": "Gcell-updates/s + % HBM roofline (Hydro FP32 vs FP64); fwd+adjoint/fwd time ratio",
        "value": r32["value"],
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": r32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: Kelvin-Helmholtz state, seeded vy noise (splitmix64)",
        "config": {"workload": "hydro2d-kh-%dx%d-fp32" % (nx, ny), "nx": nx, "ny": ny,
                   "gamma": 1.4, "cfl": 0.4, "boundary": "periodic",
                   "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "achieved": r32["achieved_GBps"], "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": r32["achieved_GBps"] / HBM_PEAK_GBPS,
                     "traffic": None,
                     "kernel": "%s_proceed_pipelined_l%d" % (m32.name, r32["dominant_launch"]),
                     "algorithmic_bytes_per_cell": 32,
                     "avg_launch_ms": r32["dominant_avg_ms"]},
        "launch_ms_per_step": r32["launch_ms_per_step"],
        "step_kernel": "proceed_pipelined",
        "unfused": {"kernel": "proceed", "value": r32u["value"],
                    "ms_per_step": r32u["ms_per_step"],
                    "launch_ms_per_step": r32u["launch_ms_per_step"]},
    }
    # the rate with the state crossing PCIe: the four fields written from host memory
    # before the run and read back after it (the C ABI also takes host buffers)
    from paraiso_amd import inputs as _inp
    host = _inp.kh_init(nx, ny, 1, dtype=np.float32)
    m32.sync()
    t0 = time.perf_counter()
    for name, arr in zip(FIELDS, host):
        m32.set(name, arr)
    t_in = time.perf_counter() - t0
    t0 = time.perf_counter()
    back = [m32.get(f) for f in FIELDS]
    t_out = time.perf_counter() - t0
    wall = r32["ms_per_step"] * K * 1e-3
    out["pcie_inclusive"] = {"value": nx * ny * K / (wall + t_in + t_out) / 1e9,
                             "unit": "Gcell-updates/s", "host_to_device_ms": t_in * 1e3,
                             "device_to_host_ms": t_out * 1e3,
                             "bytes_each_way": int(sum(a.nbytes for a in back)),
                             "note": "the %d timed steps plus writing the four fields from "
                                     "pageable host memory and reading them back" % K}
    roof = out["roofline"]
    roof["hbm"] = {"achieved": roof["achieved"], "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                   "frac": roof["frac"], "algorithmic_bytes_per_cell": 32}
    if (nx, ny) == (2048, 2048):
        traffic, valu = profile_figures("hydro_f32", roof["kernel"])
        roof["traffic"] = traffic
        if traffic is not None:
            roof["traffic_source"] = ("%s: FETCH_SIZE and WRITE_SIZE per dispatch, scaled by a "
                                      "1 GiB copy's readings" % PROFILE)
        if valu is not None:
            # the kernel is bound by VALU issue, not HBM: the roofline is the lane-instruction
            # rate against the CUs' issue rate, with the HBM fraction kept beside it
            t = r32["dominant_avg_ms"] * 1e-3
            lanes = valu["wave_instructions"] * 64
            ops = useful_ops_per_cell("hydro_f32", "proceed_pipelined")
            roof.update({"bound": "valu", "achieved": lanes / t / 1e12,
                         "peak": LANE_VALU_PEAK / 1e12,
                         "unit": "T lane-VALU instructions/s",
                         "frac": lanes / t / LANE_VALU_PEAK,
                         "lane_valu_per_cell": lanes / (nx * ny),
                         "useful_ops_per_cell": ops,
                         "useful_fp_fraction": ops["fp_arith"] / (lanes / (nx * ny)),
                         "instructions_source": "%s: SQ_INSTS_VALU per dispatch (instruction "
                                                "counts of this build, whose build.key the "
                                                "profile records); the time is this run's"
                                                % PROFILE})
            if valu["active_quad_cycles"] is not None:
                # issue slots actually used (a transcendental holds its SIMD for two), over
                # the slots the CUs offer during the live launch time
                roof["busy"] = valu["active_quad_cycles"] / (VALU_PEAK * t)
            if valu.get("mix"):
                roof["mix_per_cell"] = {k: v * 64 / (nx * ny) for k, v in valu["mix"].items()}
    if not args.no_fp64:
        m64, r64 = time_machine("hydro_f64", nx, ny, K, W, 1, np.float64, 64)
        out["fp64"] = {"value": r64["value"], "ms_per_step": r64["ms_per_step"],
                       "achieved_GBps": r64["achieved_GBps"],
                       "frac": r64["achieved_GBps"] / HBM_PEAK_GBPS,
                       "algorithmic_bytes_per_cell": 64}
        acc = accuracy(nx, ny)
        out["fp32_vs_fp64"] = acc
        out["fp32_vs_fp64_rel_l1_conserved"] = acc["rel_l1_conserved"]
        out["fp32_over_fp64_speedup"] = r32["value"] / r64["value"]
    out["hbm_copy_GBps"] = hbm_copy_gbps()
    extras = [] if args.no_extras else [x for x in args.extras.split(",") if x]
    runners = {"batched": bench_batched, "adjoint": bench_adjoint, "wave3d": bench_wave,
               "life": bench_life}
    for x in extras:
        if x not in runners:
            print("unknown extra measurement %r" % x, file=sys.stderr)
            sys.exit(2)
        out[x] = runners[x]()
    out["fwd_adj_over_fwd"] = out["adjoint"]["fwd_adj_over_fwd"] if "adjoint" in out else None
    out["cpu_baseline"] = None if args.no_cpu else cpu_baseline(nx, ny)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
