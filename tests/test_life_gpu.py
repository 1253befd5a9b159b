"""Life on the GPU against the CPU oracle: bit-exact cells, population, generation.

The oracle is the PlanTrans restatement (oracle/cpp_emitter.py) of
examples/Life/Generator.hs on Cyclic boundaries at the same grid size.
"""
import numpy as np
import pytest

from helpers import oracle_machine
from paraiso_amd import inputs, variants
from paraiso_amd.schemes.life import life_om

pytestmark = pytest.mark.gpu


def run_both(nx, ny, gens, seed, variant="life_i32"):
    print("life case %s nx=%d ny=%d gens=%d seed=%d" % (variant, nx, ny, gens, seed), flush=True)
    cells = inputs.life_init((ny, nx), seed)
    gpu = variants.machine(variant, (nx, ny))
    gpu.run("init")
    gpu.set("cell", cells)
    ora = oracle_machine(life_om, (nx, ny))
    ora.run("init")
    ora.set("cell", cells)
    gpu.run("proceed", gens)
    ora.run("proceed", gens)
    return gpu, ora


@pytest.mark.parametrize("variant", ["life_i32", "life_i32_march"])
@pytest.mark.parametrize("nx,ny,gens,seed", [(64, 64, 50, 1), (100, 37, 30, 2),
                                              (8, 8, 20, 3), (256, 128, 200, 4)])
def test_life_matches_oracle(require_gpu, nx, ny, gens, seed, variant):
    gpu, ora = run_both(nx, ny, gens, seed, variant)
    assert np.array_equal(gpu.get("cell"), ora.get("cell"))
    assert gpu.get("population") == ora.get("population")
    assert gpu.get("generation") == ora.get("generation") == gens


def test_life_baseline_size_bit_exact(require_gpu):
    """The benchmark configuration: 1024x1024, Bernoulli(0.5), 1000 generations."""
    gpu, ora = run_both(1024, 1024, 1000, 12345)
    g, o = gpu.get("cell"), ora.get("cell")
    assert np.array_equal(g, o)
    assert gpu.get("population") == ora.get("population") == int(o.sum())


def test_life_graph_replay_matches_eager(require_gpu):
    cells = inputs.life_init((96, 96), 7)
    a = variants.machine("life_i32", (96, 96))
    b = variants.machine("life_i32", (96, 96))
    for m in (a, b):
        m.run("init")
        m.set("cell", cells)
    a.run("proceed", 10)
    for _ in range(3):                      # replay one graph of 3 steps, plus a single step
        b.run("proceed", 3, graph=True)
        b.sync()
    b.run("proceed", 1, graph=True)
    a.run("proceed", 0)
    assert np.array_equal(a.get("cell"), b.get("cell"))
    assert a.get("population") == b.get("population")
    assert a.get("generation") == b.get("generation") == 10


@pytest.mark.parametrize("nx,ny,gens,k", [(1024, 1024, 1000, 4), (100, 37, 40, 4),
                                          (1024, 1024, 1000, 8), (100, 37, 40, 8),
                                          (9, 130, 24, 8)])
def test_life_several_generations_per_launch(require_gpu, nx, ny, gens, k):
    """proceed_x<k> (k generations per launch, halo k) stores what k x proceed stores."""
    print("life x%d nx=%d ny=%d gens=%d" % (k, nx, ny, gens), flush=True)
    cells = inputs.life_init((ny, nx), 77)
    gpu = variants.machine("life_i32", (nx, ny))
    gpu.run("init")
    gpu.set("cell", cells)
    gpu.run("proceed_x%d" % k, gens // k, graph=True)
    ora = oracle_machine(life_om, (nx, ny))
    ora.run("init")
    ora.set("cell", cells)
    ora.run("proceed", gens)
    assert np.array_equal(gpu.get("cell"), ora.get("cell"))
    assert gpu.get("population") == ora.get("population")
    assert gpu.get("generation") == ora.get("generation") == gens
