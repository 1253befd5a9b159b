"""Whole-machine state files (pr_state_save / pr_state_load, the paper's save-points,
paper/om.tex:679-703): a run interrupted by save and resumed from the file in a fresh
machine continues bit for bit; a file from another machine or size is rejected and
leaves the machine untouched."""
import numpy as np
import pytest

from paraiso_amd import inputs, variants
from paraiso_amd.machine import PrError

pytestmark = pytest.mark.gpu
F = ("density", "velocity0", "velocity1", "pressure")


def hydro(nx, ny, seed=3):
    m = variants.machine("hydro_f64", (nx, ny))
    for k, v in inputs.hydro_scalars(nx, ny).items():
        m.set(k, v)
    for name, arr in zip(F, inputs.kh_init(nx, ny, seed)):
        m.set(name, arr)
    return m


def test_save_resume_bit_exact(require_gpu, tmp_path):
    nx, ny = 72, 40
    path = str(tmp_path / "hydro.state")
    a = hydro(nx, ny)
    a.run("proceed_prime")
    a.run("proceed_pipelined", 6)
    a.save(path)
    a.run("proceed_pipelined", 5)
    b = variants.machine("hydro_f64", (nx, ny))
    b.load(path)
    b.run("proceed_pipelined", 5)
    assert a.get("time") == b.get("time")
    for f in F:
        assert np.array_equal(a.get(f), b.get(f)), f


def test_mismatched_state_is_rejected(require_gpu, tmp_path):
    path = str(tmp_path / "life.state")
    life = variants.machine("life_i32", (32, 32))
    life.run("init")
    life.save(path)
    other = hydro(32, 32)
    before = other.get("density")
    with pytest.raises(PrError):
        other.load(path)                       # another machine
    assert np.array_equal(other.get("density"), before)
    bigger = variants.machine("life_i32", (64, 32))
    with pytest.raises(PrError):
        bigger.load(path)                      # another size
