"""Batched Hydro: uneven grids advanced by the same launches (one graph per
sequence of steps) must equal the single-grid path grid by grid, bit for bit
(the batched kernels share the single-grid kernels' code)."""
import numpy as np
import pytest

from paraiso_amd import inputs, variants

pytestmark = pytest.mark.gpu
F = ("density", "velocity0", "velocity1", "pressure")
SIZES = [(64, 48), (100, 37), (33, 90), (128, 64), (70, 70), (9, 16)]


def init(m, nx, ny, seed, dt):
    for k, v in inputs.hydro_scalars(nx, ny).items():
        m.set(k, v)
    for name, arr in zip(F, inputs.kh_init(nx, ny, seed, dtype=dt)):
        m.set(name, arr)


@pytest.mark.parametrize("variant,dt", [("hydro_f32", np.float32), ("hydro_f64", np.float64)])
def test_batch_equals_single_grids(require_gpu, variant, dt):
    steps = 12
    b = variants.batch(variant, SIZES)
    for g, (nx, ny) in enumerate(SIZES):
        init(b.grids[g], nx, ny, 100 + g, dt)
    b.run("proceed", 5)
    b.run("proceed", 7, graph=True)
    b.sync()
    for g, (nx, ny) in enumerate(SIZES):
        print("batch grid %d %dx%d" % (g, nx, ny), flush=True)
        m = variants.machine(variant, (nx, ny))
        init(m, nx, ny, 100 + g, dt)
        m.run("proceed", steps)
        assert b.grids[g].get("time") == m.get("time")
        for f in F:
            assert np.array_equal(b.grids[g].get(f), m.get(f)), (g, f)


@pytest.mark.parametrize("variant", ["hydro_f32", "hydro_f32_batch"])
def test_batch_pipelined_equals_single_grids(require_gpu, variant):
    """The carried-CFL step batched over uneven grids, replayed as one graph."""
    steps = 9
    b = variants.batch(variant, SIZES)
    for g, (nx, ny) in enumerate(SIZES):
        init(b.grids[g], nx, ny, 200 + g, np.float32)
    b.run("proceed_prime")
    b.run("proceed_pipelined", steps, graph=True)
    b.sync()
    for g, (nx, ny) in enumerate(SIZES):
        print("batch pipelined grid %d %dx%d" % (g, nx, ny), flush=True)
        m = variants.machine(variant, (nx, ny))
        init(m, nx, ny, 200 + g, np.float32)
        m.run("proceed_prime")
        m.run("proceed_pipelined", steps)
        assert b.grids[g].get("time") == m.get("time")
        for f in F:
            assert np.array_equal(b.grids[g].get(f), m.get(f)), (g, f)


def test_batch_config_sizes_bit_identical(require_gpu):
    """32 grids with the benchmark's size law (sides log-uniform on [64, 1024], drawn
    independently for nx and ny, examples/Hydro/HydroMain.hs run per grid), 20 carried-CFL
    steps replayed as one graph: each grid equals its own single-grid run bit for bit."""
    steps, seed = 20, 2024
    sizes = [tuple(int(x) for x in s) for s in inputs.batched_sizes(32, seed)]
    assert min(min(s) for s in sizes) >= 64 and max(max(s) for s in sizes) <= 1024
    print("batch of %d grids, %d cells" % (len(sizes), sum(a * b for a, b in sizes)), flush=True)
    b = variants.batch("hydro_f32_batch", sizes)
    for g, (nx, ny) in enumerate(sizes):
        init(b.grids[g], nx, ny, seed + g, np.float32)
    b.run("proceed_prime")
    b.run("proceed_pipelined", steps, graph=True)
    b.sync()
    for g, (nx, ny) in enumerate(sizes):
        print("config-size grid %d %dx%d" % (g, nx, ny), flush=True)
        m = variants.machine("hydro_f32_batch", (nx, ny))
        init(m, nx, ny, seed + g, np.float32)
        m.run("proceed_prime")
        m.run("proceed_pipelined", steps)
        assert b.grids[g].get("time") == m.get("time")
        for f in F:
            assert np.array_equal(b.grids[g].get(f), m.get(f)), (g, f)
        m.close()
