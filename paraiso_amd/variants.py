"""The machines this package builds ahead of time, one shared library each.

Grid sizes are run-time parameters of a machine, so one library serves every
size of its program.  The ratio of specific heats and the dt mode are baked
into the OM of a Hydro program, as in the reference.
"""
from __future__ import annotations

from fractions import Fraction

from . import annotation as A
from .generator import Language, Setup, build_machine
from .om.dyn import CType

CYC2 = (A.Condition.Cyclic,) * 2
CYC3 = (A.Condition.Cyclic,) * 3
CYC1 = (A.Condition.Cyclic,)
OPEN2 = (A.Condition.Open,) * 2


def _life():
    """Life with ``proceed_x4`` and ``proceed_x8``: four or eight generations per launch
    (temporal blocking)."""
    from .schemes.life import life_om
    from .transform import unroll
    return unroll(unroll(life_om(), "proceed", 4), "proceed", 8)


def _life_plain():
    from .schemes.life import life_om
    return life_om()


def _life_legacy():
    from .schemes.life import life_om_legacy
    return life_om_legacy()


def _hydro_legacy():
    from .schemes.hydro import hydro_om_legacy
    return hydro_om_legacy()


def _hydro(real, gamma=Fraction(7, 5), dt_mode="cfl"):
    """Hydro with, beside ``proceed``, ``proceed_prime`` + ``proceed_pipelined``: the CFL
    minimum of the next step computed in the store phase of this one (one grid pass)."""
    def mk():
        from .schemes.hydro import hydro_om
        from .transform import pipeline_reduce
        om = hydro_om(real, gamma, dt_mode)
        return pipeline_reduce(om, "proceed") if dt_mode == "cfl" else om
    return mk


def _wave(ndim=3, real=CType.Float):
    def mk():
        from .schemes.linear_wave import linear_wave_om
        return linear_wave_om(ndim, real)
    return mk


HYDRO_FIELDS = ["density", "velocity0", "velocity1", "pressure"]


def _hydro_adj(real):
    def mk():
        from .adjoint import add_adjoint
        from .schemes.hydro import hydro_om
        return add_adjoint(hydro_om(real, Fraction(7, 5), "fixed"), "proceed", HYDRO_FIELDS)
    return mk


STAGE_FIELDS = ["stage_" + f for f in HYDRO_FIELDS]


def _hydro_adj_staged(real):
    """The fixed-dt step split into its two stages (``proceed_stage1`` stores the half-step
    state, ``proceed_stage2`` finishes the step from it), with one adjoint kernel per stage:
    the checkpoint driver replays both stages and runs ``proceed_stage2_adjoint`` then
    ``proceed_stage1_adjoint``, so no adjoint launch recomputes stage 1 under the halo that
    stage 2 needs."""
    def mk():
        from .adjoint import add_adjoint
        from .schemes.hydro import hydro_om
        om = hydro_om(real, Fraction(7, 5), "fixed", stages=True)
        om = add_adjoint(om, "proceed_stage2", HYDRO_FIELDS + STAGE_FIELDS,
                         zero_in=STAGE_FIELDS)
        return add_adjoint(om, "proceed_stage1", HYDRO_FIELDS + STAGE_FIELDS)
    return mk


def _wave_adj(ndim, real):
    """LinearWave with its adjoint, and both fused two steps per launch (``proceed_x2``,
    ``proceed_adjoint_x2``; more steps per launch need deeper halos than a streaming tile's
    LDS holds, and lose to two)."""
    def mk():
        from .adjoint import add_adjoint
        from .schemes.linear_wave import linear_wave_om
        from .transform import unroll
        om = add_adjoint(linear_wave_om(ndim, real), "proceed", ["fieldF", "fieldG"])
        return unroll(unroll(om, "proceed", 2), "proceed_adjoint", 2)
    return mk


def _optable(real):
    def mk():
        from .schemes.optable import optable_om
        return optable_om(real)
    return mk


# name -> (OM factory, setup, precise)
VARIANTS = {
    "life_i32": (_life, Setup((64, 64), Language.HIP, boundary=CYC2), True),
    "hydro_f32": (_hydro(CType.Float), Setup((64, 64), Language.HIP, boundary=CYC2, batched=True),
                  False),
    # the headline machine's launch form (tile, work-items, resident tile loop), correctly
    # rounded: checked bit for bit against the oracle
    "hydro_f32_precise": (_hydro(CType.Float), Setup((64, 64), Language.HIP, boundary=CYC2,
                                                     tile=(64, 32), threads=1024, persist=True),
                          True),
    # the same machine, tuned for batches of small uneven grids
    "hydro_f32_batch": (_hydro(CType.Float), Setup((64, 64), Language.HIP, boundary=CYC2,
                                                   batched=True), False),
    "hydro_f64": (_hydro(CType.Double), Setup((64, 64), Language.HIP, boundary=CYC2, batched=True),
                  True),
    "hydro_adj_f32": (_hydro_adj(CType.Float), Setup((64, 64), Language.HIP, boundary=CYC2,
                                                     tile=(32, 8)), True),
    # FP32 bulk arithmetic with the hardware reciprocal and square root in the adjoint too
    "hydro_adj_f32_fast": (_hydro_adj(CType.Float), Setup((64, 64), Language.HIP, boundary=CYC2,
                                                          tile=(32, 8)), False),
    # the fixed-dt step in FP32 (fast arithmetic) at the headline's launch form
    "hydro_fixed_f32": (_hydro(CType.Float, dt_mode="fixed"),
                        Setup((64, 64), Language.HIP, boundary=CYC2, tile=(64, 32), threads=1024,
                              persist=True), False),
    # the plain FP64 fixed-dt step at its fastest launch form (the comparison for the adjoint)
    "hydro_fixed_f64": (_hydro(CType.Double, dt_mode="fixed"),
                        Setup((64, 64), Language.HIP, boundary=CYC2, tile=(32, 32), threads=1024,
                              persist=True), True),
    "hydro_adj_f64": (_hydro_adj(CType.Double), Setup((64, 64), Language.HIP, boundary=CYC2,
                                                      tile=(32, 8)), True),
    # the adjoint per stage (see _hydro_adj_staged), FP32 fast and FP64 correctly rounded
    "hydro_adj2_f32_fast": (_hydro_adj_staged(CType.Float),
                            Setup((64, 64), Language.HIP, boundary=CYC2, tile=(32, 8)), False),
    "hydro_adj2_f64": (_hydro_adj_staged(CType.Double),
                       Setup((64, 64), Language.HIP, boundary=CYC2, tile=(32, 8)), True),
    # the same with the adjoint kernels' shared expansions joined into earlier phases
    # (correctly rounded, so bit-identical to the above)
    "hydro_adj2_f64_join": (_hydro_adj_staged(CType.Double),
                            Setup((64, 64), Language.HIP, boundary=CYC2, tile=(32, 16),
                                  threads=512,
                                  kernel_forms=(("proceed_stage1_adjoint", {"join": True}),
                                                ("proceed_stage2_adjoint", {"join": True}))),
                            True),
    "wave3d_f32": (_wave_adj(3, CType.Float), Setup((64, 64, 64), Language.HIP, boundary=CYC3), False),
    "wave3d_f64": (_wave_adj(3, CType.Double), Setup((64, 64, 64), Language.HIP, boundary=CYC3), True),
    # streaming along the last axis (rings of planes in LDS), checked bit for bit
    "life_i32_march": (_life, Setup((64, 64), Language.HIP, boundary=CYC2, tile=(64, 32),
                                    march=4), True),
    "hydro_f64_march": (_hydro(CType.Double), Setup((64, 64), Language.HIP, boundary=CYC2,
                                                    tile=(64, 32), march=4), True),
    # a tile whose load and annotated-value tiles alone overflow LDS: the scheduler reads the
    # loads from memory and evaluates some annotated values in registers, checked bit for bit
    "hydro_f64_shed": (_hydro(CType.Double), Setup((64, 64), Language.HIP, boundary=CYC2,
                                                   tile=(64, 32), threads=1024), True),
    "wave3d_f64_march": (_wave_adj(3, CType.Double), Setup((64, 64, 64), Language.HIP,
                                                           boundary=CYC3, tile=(32, 8, 32),
                                                           march=1), True),
    # the reference's own LinearWave is 1-D (Convergence.hs:25-54): 1-D and 2-D machines of
    # the same program, checked bit for bit
    "wave1d_f64": (_wave_adj(1, CType.Double), Setup((256,), Language.HIP, boundary=CYC1,
                                                     tile=(256,), threads=256), True),
    "wave2d_f64": (_wave_adj(2, CType.Double), Setup((64, 64), Language.HIP, boundary=CYC2,
                                                     tile=(32, 16), threads=512), True),
    # block tiles with the resident tile loop, whose load phase computes the half step f1
    # from the next tile's prefetched fields, checked bit for bit
    "wave3d_f64_persist": (_wave_adj(3, CType.Double), Setup((64, 64, 64), Language.HIP,
                                                             boundary=CYC3, tile=(16, 8, 8),
                                                             threads=512, persist=True), True),
    # every operator the emitters lower, one per static (checked against the oracle)
    "optable_f32": (_optable(CType.Float), Setup((64, 48), Language.HIP, boundary=CYC2), True),
    "optable_f32_fast": (_optable(CType.Float), Setup((64, 48), Language.HIP, boundary=CYC2),
                         False),
    "optable_f64": (_optable(CType.Double), Setup((64, 48), Language.HIP, boundary=CYC2), True),
    "optable_i32": (_optable(CType.Int), Setup((64, 48), Language.HIP, boundary=CYC2), True),
    # Open (non-periodic) boundaries: ghost margins, as the reference's shipped examples
    "life_open_i32": (_life_plain, Setup((128, 128), Language.HIP, boundary=OPEN2), True),
    "life_legacy_open_i32": (_life_legacy, Setup((128, 128), Language.HIP, boundary=OPEN2), True),
    "hydro_legacy_open_f32": (_hydro_legacy, Setup((1024, 1024), Language.HIP, boundary=OPEN2),
                              True),
}


def setup_of(name: str):
    """The variant's setup, with the tile the autotuner recorded (paraiso_amd/tuned.json)."""
    from .autotune import tuned
    mk, setup, precise = VARIANTS[name]
    rec = tuned().get(name)
    if rec is not None and "tile" in rec:
        setup = setup.with_(tile=tuple(rec["tile"]), threads=int(rec.get("threads", 256)),
                            mat_gain=int(rec.get("mat_gain", 48)), march=int(rec.get("march", 0)),
                            persist=bool(rec.get("persist", setup.persist)),
                            join=bool(rec.get("join", setup.join)))
    if rec is not None and "kernel_forms" in rec:
        forms = []
        for k, f in sorted(rec["kernel_forms"].items()):
            form = {}
            if "tile" in f:
                form.update({"tile": tuple(f["tile"]), "threads": int(f["threads"]),
                             "mat_gain": int(f.get("mat_gain", 48)),
                             "march": int(f.get("march", 0)),
                             "persist": bool(f.get("persist", False)),
                             "join": bool(f.get("join", False))})
            # the per-node Manifest/Delayed choices the autotuner kept
            for key in ("manifest", "delayed"):
                if f.get(key):
                    form[key] = tuple(int(v) for v in f[key])
            forms.append((k, form))
        setup = setup.with_(kernel_forms=tuple(forms))
    return setup


def build(name: str, force: bool = False) -> str:
    mk, _, precise = VARIANTS[name]
    return build_machine(setup_of(name), mk(), name, precise=precise, force=force)


def build_all(force: bool = False, jobs: int = 4) -> dict:
    """Build every variant in ``jobs`` worker processes (scheduling is Python, compiling is
    hipcc; both run in parallel across variants).  The slowest variants start first."""
    from concurrent.futures import ProcessPoolExecutor
    order = sorted(VARIANTS, key=lambda n: ("adj" not in n, n))
    with ProcessPoolExecutor(max_workers=max(1, jobs)) as ex:
        futs = {n: ex.submit(build, n, force) for n in order}
        return {n: f.result() for n, f in futs.items()}


def library(name: str) -> str:
    import os
    from .generator import GEN_DIR
    return os.path.join(GEN_DIR, name, "lib%s.so" % name)


def machine(name: str, local_size):
    from .machine import Machine
    return Machine(library(name), local_size)


def batch(name: str, sizes):
    from .machine import Batch
    return Batch(library(name), sizes)
