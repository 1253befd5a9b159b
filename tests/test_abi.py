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
    assert d.abi == 4
    mk, setup, _ = variants.VARIANTS[name]
    om = mk()
    user = [d.statics[i].name.decode() for i in range(d.n_static) if not d.statics[i].flags & 1]
    assert user == [s.name for s in om.setup.static_values]
    assert [d.kernels[i].name.decode() for i in range(d.n_kernel)] == [k.name for k in om.kernels]


def _kernel_statics(name, kernel, field):
    lib = ctypes.CDLL(variants.library(name))
    lib.pm_describe.restype = ctypes.POINTER(PrMachineDesc)
    d = lib.pm_describe().contents
    names = [d.statics[i].name.decode() for i in range(d.n_static)]
    k = next(d.kernels[i] for i in range(d.n_kernel) if d.kernels[i].name.decode() == kernel)
    n = getattr(k, "n_" + field)
    return {names[getattr(k, field)[i]] for i in range(n)}


def test_descriptors_list_what_kernels_read_and_write():
    """The adjoint driver skips the forward replay when the adjoint kernel reads nothing the
    forward kernel writes: true of the linear wave, not of Hydro."""
    wave_adj = _kernel_statics("wave3d_f32", "proceed_adjoint", "loaded")
    wave_fwd = _kernel_statics("wave3d_f32", "proceed", "written")
    assert {"fieldF", "fieldG"} <= wave_fwd and "energy" in wave_fwd
    assert not (wave_adj & wave_fwd) and {"adj_fieldF", "adj_fieldG"} <= wave_adj
    hyd_adj = _kernel_statics("hydro_adj_f32", "proceed_adjoint", "loaded")
    hyd_fwd = _kernel_statics("hydro_adj_f32", "proceed", "written")
    assert {"density", "pressure"} <= (hyd_adj & hyd_fwd)
