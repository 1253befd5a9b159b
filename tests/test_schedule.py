"""The HIP launch scheduler (paraiso_amd/generator/hip/schedule.py) on the CPU.

A launch starts from LDS tiles of its loads and of the values the program
annotates Manifest.  When those alone overflow the 160 KiB of LDS a workgroup
may use (large FP64 tiles, or a launch behind a cut that reads many scratch
arrays), the scheduler reads the least reused loads straight from memory and
evaluates the least reused annotated values in registers instead of failing.
The GPU tests check such a machine (``hydro_f64_shed``) bit for bit against
the oracle."""
from fractions import Fraction

from paraiso_amd import annotation as A
from paraiso_amd.generator.hip.ir import lower_kernel
from paraiso_amd.generator.hip.schedule import LDS_MAX, lds_budget, schedule_kernel
from paraiso_amd.generator.native import Language, Setup
from paraiso_amd.generator.plan import translate
from paraiso_amd.om.dyn import CType
from paraiso_amd.schemes.hydro import hydro_om
from paraiso_amd.transform import pipeline_reduce

CYC2 = (A.Condition.Cyclic,) * 2


def _pipelined_f64():
    om = pipeline_reduce(hydro_om(CType.Double, Fraction(7, 5)), "proceed")
    plan = translate(Setup((64, 64), Language.HIP, boundary=CYC2), om)
    ki = [k.name for k in plan.om.kernels].index("proceed_pipelined")
    return lower_kernel(plan.om, ki, (True, True))


def test_overflowing_start_tiles_are_shed():
    k = _pipelined_f64()
    for tile, threads, gain in (((64, 64), 1024, 24), ((64, 32), 1024, 48)):
        s = schedule_kernel(k, tile, lds_budget(threads), gain)
        assert all(la.lds_bytes <= LDS_MAX for la in s.launches), tile
        # the four primitive fields are read from memory, not staged
        assert any(len(la.loads_global) == 4 for la in s.launches), tile


def test_fitting_tiles_shed_nothing(monkeypatch, capsys):
    """Shedding only applies when the starting tiles overflow: the FP64 tile of 32x16 cells
    and 512 work-items sheds nothing."""
    from paraiso_amd.generator.hip import schedule as S
    monkeypatch.setattr(S, "SCHED_DEBUG", True)
    k = _pipelined_f64()
    schedule_kernel(k, (32, 16), lds_budget(512), 48)
    out = capsys.readouterr().out
    assert "schedule L" in out and "shed" not in out


def test_reduction_scratch_is_left_free():
    """A kernel declares one 8-byte slot per wave for each Reduce it finishes beside the
    materialised values, and the two together must fit the CU's LDS: the FP64 step at 32x32
    cells and 1024 work-items once filled LDS to the last byte without counting them."""
    from paraiso_amd.generator.hip.schedule import RED_SCRATCH
    k = _pipelined_f64()
    s = schedule_kernel(k, (32, 32), lds_budget(1024), 48)
    for la in s.launches:
        assert la.lds_bytes + RED_SCRATCH * len(la.reduces) <= LDS_MAX


def test_per_node_allocation_choices_are_honoured():
    """The autotuner's per-node Manifest/Delayed choice (DecideAllocation.hs:41-43): a value
    forced Delayed is expanded in registers, one forced Manifest is kept in LDS, and the
    other values keep the threshold's choice (FP32 step at the headline's launch form).  A
    field may leave LDS: a value computed point by point from it is computed while the field
    is loaded and kept in its place."""
    om = pipeline_reduce(hydro_om(CType.Float, Fraction(7, 5)), "proceed")
    plan = translate(Setup((64, 64), Language.HIP, boundary=CYC2), om)
    ki = [k.name for k in plan.om.kernels].index("proceed_pipelined")
    k = lower_kernel(plan.om, ki, (True, True), exact=True)
    base = schedule_kernel(k, (64, 32), lds_budget(1024), 48)
    assert len(base.launches) == 1
    la = base.launches[0]
    mat = sorted(v for g, v, m in la.choices if m)
    exp = [v for g, v, m in la.choices if not m]
    d, m = mat[0], exp[0]
    s = schedule_kernel(k, (64, 32), lds_budget(1024), 48, force_manifest=[m],
                        force_delayed=[d])
    got = s.launches[0].mat
    assert m in got and d not in got
    computed = lambda vs: set(v for v in vs if k.nodes[v].op != "load")
    assert computed(got) - {m} == computed(la.mat) - {d}


def test_autotuner_flips_are_nearest_the_threshold():
    from paraiso_amd import autotune
    flips = autotune.alloc_candidates("hydro_f32", "proceed_pipelined", 4)
    assert len(flips) == 4 and {f[0] for f in flips} <= {"manifest", "delayed"}
    assert len(set(v for _, v in flips)) == 4


def test_pointwise_values_join_the_load_phase():
    """A value computed point by point from loaded fields is computed while they are loaded.
    In the two-step LinearWave launch, f1 = f0 + dt g0 (Convergence.hs:44) joins phase 0 and
    is kept in LDS in place of f0, which nothing else reads, although it saves fewer
    operations than the threshold asks of a value that adds a phase; g0 stays, as the update
    of g reads it; the streaming kernel stores f1 from the prefetched registers, reading no
    field at an offset in its load phase; and forced Delayed, f1 is expanded again."""
    import re
    from paraiso_amd.generator.hip.emit import MachineEmitter
    from paraiso_amd.schemes.linear_wave import linear_wave_om
    from paraiso_amd.transform import unroll
    om = unroll(linear_wave_om(3, CType.Float), "proceed", 2)
    plan = translate(Setup((64, 64, 64), Language.HIP, boundary=(A.Condition.Cyclic,) * 3), om)
    ki = [k.name for k in plan.om.kernels].index("proceed_x2")
    k = lower_kernel(plan.om, ki, (True,) * 3, exact=True)
    loads = [n.id for n in k.nodes if n.op == "load" and n.realm.name == "Array"]
    f0, g0 = loads[0], loads[1]
    f1 = next(n.id for n in k.nodes if n.op == "arith" and f0 in n.args)
    tile = (64, 16, 256)
    delayed = schedule_kernel(k, tile, lds_budget(1024), 48, march=2, force_delayed=[f1])
    assert f1 not in delayed.launches[0].mat and f0 in delayed.launches[0].mat
    s = schedule_kernel(k, tile, lds_budget(1024), 48, march=2)
    la = s.launches[0]
    ph0 = la.phases[0]
    assert ph0.index == 0 and sorted(ph0.nodes) == sorted([g0, f1])
    assert f0 not in la.mat and f1 in la.lds
    # f1 is stored once per point, and read at the Laplacian's seven points
    assert len(la.lc[f1][1]) == 7
    em = MachineEmitter("W", plan.om.setup.static_values, 3, [s], tile, fast_fp32=True,
                        threads=1024, march=2)
    src, _ = em.emit_launch(0, s, la)
    ring = src[src.index("into their rings"):]
    ring = ring[:ring.index("__syncthreads")]
    assert re.search(r"s%d\[[^\]]*\] = n%d_0_0_0;" % (f1, f1), ring)
    assert "a.ld" not in ring and "pf%d_0" % f0 in ring


def test_resident_loop_is_a_candidate_dimension():
    """A candidate may fix the resident tile loop (a fifth element); without it the variant's
    own choice stands and the candidate keeps its earlier name."""
    from paraiso_amd import autotune
    assert autotune._unpack(((64, 32), 1024)) == ((64, 32), 1024, 48, 0, None, False)
    assert autotune._unpack(((64, 32), 1024, 48, 0, True))[4] is True
    assert autotune.candidate_name("v", (64, 32), 1024) == "v__t64x32_n1024"
    assert autotune.candidate_name("v", (64, 32), 1024, 48, 0, True) == "v__t64x32_n1024_p"
    assert autotune.candidate_name("v", (64, 32), 1024, 48, 0, False) == "v__t64x32_n1024_np"
    for c in autotune.CANDIDATE_SETS["headline_persist"]:
        assert autotune._unpack(c)[4] is not None


def _stage2_adjoint_f32():
    from paraiso_amd import variants
    om = variants._hydro_adj_staged(CType.Float)()
    plan = translate(Setup((64, 64), Language.HIP, boundary=CYC2), om)
    ki = [k.name for k in plan.om.kernels].index("proceed_stage2_adjoint")
    return lower_kernel(plan.om, ki, (True, True), exact=True)


def test_join_computes_shared_expansions_once():
    """The stage-2 Hydro adjoint: the cell update reads each face's left- and right-state
    cotangents once, and each of them expands into the face's forward HLLC work, which the
    phase that materialises the face's other cotangents evaluates already.  With ``join``
    they are computed there, as one group, and read from LDS: the launch stays whole within
    the CU's LDS and the store phase keeps under half of its register expansion."""
    from paraiso_amd.generator.hip.schedule import OUT
    k = _stage2_adjoint_f32()
    nodes = k.nodes

    def out_evals(la):
        return sum(len(c) for v, ctxs in la.lc.items() if nodes[v].op == "arith"
                   for ctx, c in ctxs.items() if ctx == OUT)
    cut = schedule_kernel(k, (64, 8), lds_budget(512), 48)
    full = schedule_kernel(k, (64, 8), LDS_MAX, 48)
    joined = schedule_kernel(k, (64, 8), lds_budget(512), 48, join=True)
    assert len(cut.launches) == 2 and len(full.launches) == 1
    assert len(joined.launches) == 1 and joined.launches[0].lds_bytes <= LDS_MAX
    fl, jl = full.launches[0], joined.launches[0]
    assert out_evals(jl) < 0.5 * out_evals(fl)
    group = jl.mat - fl.mat
    phase = {v: ph.index for ph in jl.phases[:-1] for v in ph.nodes}
    assert len(group) >= 16
    assert all(phase[v] < jl.phases[-1].index for v in group)


def test_per_kernel_selection_records_join_and_resident_loop(tmp_path, monkeypatch):
    """select --per-kernel keeps, per OM kernel, the fastest candidate's whole form: its
    join and resident-loop choices too, and setup_of hands them to the generator."""
    import json
    from paraiso_amd import autotune, variants
    tuned = tmp_path / "tuned.json"
    tuned.write_text("{}")
    monkeypatch.setattr(autotune, "TUNED", str(tuned))
    res = [{"tile": [64, 8], "threads": 512, "mat_gain": 48, "march": 0, "persist": None,
            "join": False, "per_kernel_ms": {"a": 2.0, "b": 1.0}},
           {"tile": [32, 16], "threads": 512, "mat_gain": 48, "march": 0, "persist": True,
            "join": True, "per_kernel_ms": {"a": 1.0, "b": 3.0}}]
    runs = tmp_path / "runs.json"
    runs.write_text(json.dumps({"hydro_adj2_f32_fast": {"size": [512, 512], "steps": 40,
                                                        "results": res}}))
    t = autotune.select_per_kernel(str(runs))
    forms = t["hydro_adj2_f32_fast"]["kernel_forms"]
    assert forms["a"]["join"] is True and forms["a"]["persist"] is True
    assert forms["a"]["tile"] == [32, 16]
    assert "join" not in forms["b"] and "persist" not in forms["b"]
    assert set(t["hydro_adj2_f32_fast"]["candidates_per_kernel"]["a"]) == {
        "64x8_n512_g48_m0", "32x16_n512_g48_m0_p_j"}
    monkeypatch.setattr(autotune, "tuned", lambda: json.loads(tuned.read_text()))
    setup = variants.setup_of("hydro_adj2_f32_fast")
    fa, fb = setup.form_of("a"), setup.form_of("b")
    assert fa["join"] is True and fa["persist"] is True and fa["tile"] == (32, 16)
    assert fb["join"] is False and fb["tile"] == (64, 8)
