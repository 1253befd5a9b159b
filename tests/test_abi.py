"""The C ABI: every function include/paraiso_hip.h declares is exported by the
runtime library, and every machine library exports pm_describe with the
statics and kernels of its program.  No GPU call is made here."""
import ctypes
import os
import re
import subprocess

import pytest

from paraiso_amd import variants
from paraiso_amd.generator import INCLUDE_DIR, runtime_library
from paraiso_amd.machine import PrMachineDesc

HEADER = os.path.join(INCLUDE_DIR, "paraiso_hip.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pr_[a-z_]+)\s*\(", txt)))


def exported(so):
    out = subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True, text=True,
                         check=True).stdout
    return set(line.split()[-1] for line in out.splitlines() if " T " in line)


def test_runtime_exports_every_declared_function():
    so = runtime_library()
    assert os.path.exists(so), "run __graft_entry__.build() first"
    missing = set(declared_functions()) - exported(so)
    assert not missing, missing


@pytest.mark.parametrize("name", sorted(variants.VARIANTS))
def test_machine_library_describes_its_program(name):
    so = variants.library(name)
    assert os.path.exists(so), "run __graft_entry__.build() first"
    assert "pm_describe" in exported(so)
    lib = ctypes.CDLL(so)
    lib.pm_describe.restype = ctypes.POINTER(PrMachineDesc)
    d = lib.pm_describe().contents
    assert d.abi == 3
    mk, setup, _ = variants.VARIANTS[name]
    om = mk()
    user = [d.statics[i].name.decode() for i in range(d.n_static) if not d.statics[i].flags & 1]
    assert user == [s.name for s in om.setup.static_values]
    assert [d.kernels[i].name.decode() for i in range(d.n_kernel)] == [k.name for k in om.kernels]
