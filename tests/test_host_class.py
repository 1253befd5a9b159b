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
