"""The generated C++ host class (``<Name>.hpp``) drives a machine through the C ABI the
way the reference's examples drive their generated class (examples/Life/main.cpp):
tests/host/life_main.cpp constructs ``Life``, calls ``init``, sets the cells, calls
``proceed`` and reads the cells back.  Compiled on the host (g++, no HIP headers); run
on the GPU and compared with the oracle bit for bit."""
import os
import subprocess

import numpy as np
import pytest

from helpers import oracle_machine
from paraiso_amd import host_example, inputs, variants
from paraiso_amd.schemes.life import life_om


@pytest.mark.parametrize("name", ["life_i32", "hydro_f32", "wave3d_f64", "life_legacy_open_i32",
                                  "hydro_legacy_open_f32"])
def test_generated_classes_compile(name, tmp_path):
    """Every generated class (lazy host mirrors, elemental accessors with the plan's margins)
    is valid C++ against include/paraiso_hip.h."""
    from paraiso_amd.generator import INCLUDE_DIR
    from paraiso_amd.generator.hip.host import emit_host_class
    from paraiso_amd.generator.plan import translate
    mk, setup, _ = variants.VARIANTS[name]
    plan = translate(setup, mk())
    hpp = emit_host_class(plan.om, setup, (plan.lower_margin, plan.upper_margin))
    (tmp_path / "M.hpp").write_text(hpp)
    (tmp_path / "t.cpp").write_text('#include "M.hpp"\nint main() { return 0; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I" + INCLUDE_DIR,
                        "-I" + str(tmp_path), str(tmp_path / "t.cpp")], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr[:2000]
    if plan.lower_margin[0]:
        # the elemental accessor offsets local coordinates by the margins (PlanTrans.hs:140-150)
        lm, mem = plan.lower_margin, plan.memory_size()
        assert "(%d + i0) + %d * (%d + i1)" % (lm[0], mem[0], lm[1]) in hpp


def test_host_example_compiles():
    exe = host_example.build()
    assert os.access(exe, os.X_OK)
    hpp = open(os.path.join(host_example.OUT_DIR, "Life.hpp")).read()
    assert "void proceed()" in hpp and "pr_create" in hpp and "om_memory_size() const" in hpp


@pytest.mark.gpu
def test_host_example_matches_oracle(require_gpu, tmp_path):
    nx, ny = host_example.SIZE
    gens = 60
    exe = host_example.executable()
    assert os.access(exe, os.X_OK), "run __graft_entry__.build() first"
    init = inputs.life_init((ny, nx), 17)
    src, dst = tmp_path / "init.bin", tmp_path / "out.bin"
    init.astype(np.int32).tofile(src)
    print("host example Life %dx%d gens=%d" % (nx, ny, gens), flush=True)
    r = subprocess.run([exe, variants.library("life_i32"), str(src), str(gens), str(dst)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    pop, gen = (int(x) for x in r.stdout.split())
    ora = oracle_machine(life_om, (nx, ny))
    ora.run("init")
    ora.set("cell", init)
    ora.run("proceed", gens)
    out = np.fromfile(dst, dtype=np.int32).reshape(ny, nx)
    assert np.array_equal(out, ora.get("cell").reshape(ny, nx))
    assert pop == ora.get("population") and gen == ora.get("generation") == gens
