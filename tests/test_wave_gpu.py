"""Linear wave (examples-old/LinearWave/Convergence.hs restated in 3-D) on the GPU:
the oracle, the exact discrete plane wave of the leapfrog scheme, and the
adjoint's dot-product identity through the checkpoint driver."""
import numpy as np
import pytest
import torch

from helpers import oracle_machine
from paraiso_amd import variants
from paraiso_amd.adjoint import add_adjoint
from paraiso_amd.autograd import rollout
from paraiso_amd.om.dyn import CType
from paraiso_amd.schemes.linear_wave import linear_wave_om

pytestmark = pytest.mark.gpu


def params(m, c=1.0, dx=1.0, dt=0.5):
    m.set("c", c)
    m.set("dx", dx)
    m.set("dt", dt)


@pytest.mark.parametrize("variant,n", [("wave3d_f64", (16, 12, 10)),
                                       ("wave3d_f64_march", (16, 12, 10)),
                                       ("wave3d_f64_march", (40, 9, 70)),
                                       ("wave3d_f64_persist", (16, 12, 10)),
                                       # more tiles than resident workgroups: each takes
                                       # several, loading the next one's fields early
                                       ("wave3d_f64_persist", (128, 96, 112))])
def test_wave3d_f64_bit_exact_vs_oracle(require_gpu, variant, n):
    print("wave case %s %s" % (variant, n), flush=True)
    rng = np.random.default_rng(3)
    f0, g0 = rng.standard_normal(n[::-1]), rng.standard_normal(n[::-1])
    gpu = variants.machine(variant, n)
    ora = oracle_machine(lambda: add_adjoint(linear_wave_om(3, CType.Double), "proceed",
                                             ["fieldF", "fieldG"]), n)
    for m in (gpu, ora):
        params(m)
        m.set("fieldF", f0)
        m.set("fieldG", g0)
        m.run("proceed", 9)
    assert np.array_equal(gpu.get("fieldF"), ora.get("fieldF"))
    assert np.array_equal(gpu.get("fieldG"), ora.get("fieldG"))
    assert abs(gpu.get("energy") - ora.get("energy")) <= 1e-12 * abs(ora.get("energy"))


@pytest.mark.parametrize("variant,ndim,n", [("wave1d_f64", 1, (1000,)),
                                            ("wave2d_f64", 2, (100, 37))])
def test_wave_1d_2d_bit_exact_vs_oracle(require_gpu, variant, ndim, n):
    """The reference's LinearWave is 1-D (Convergence.hs:34-54, with c = 3.43, dx = 2 pi / n
    and dt = dx / c); the same program in 1-D and 2-D, one and two steps per launch, bit for
    bit against the oracle."""
    print("wave case %s %s" % (variant, n), flush=True)
    rng = np.random.default_rng(11)
    f0, g0 = rng.standard_normal(n[::-1]), rng.standard_normal(n[::-1])
    dx = 2 * np.pi / n[0]
    for kern, launches in (("proceed", 12), ("proceed_x2", 6)):
        gpu = variants.machine(variant, n)
        ora = oracle_machine(lambda: add_adjoint(linear_wave_om(ndim, CType.Double), "proceed",
                                                 ["fieldF", "fieldG"]), n)
        for m in (gpu, ora):
            params(m, 3.43, dx, dx / 3.43)
            m.set("fieldF", f0)
            m.set("fieldG", g0)
        gpu.run(kern, launches)
        ora.run("proceed", 12)
        assert np.array_equal(gpu.get("fieldF"), ora.get("fieldF"))
        assert np.array_equal(gpu.get("fieldG"), ora.get("fieldG"))
        assert abs(gpu.get("energy") - ora.get("energy")) <= 1e-12 * abs(ora.get("energy"))


def test_wave3d_exact_plane_wave(require_gpu):
    n = (32, 16, 8)
    c, dx, dt, steps = 1.0, 1.0, 0.5, 40
    k = np.array([2 * np.pi * 3 / n[0], 2 * np.pi * 1 / n[1], 2 * np.pi * 2 / n[2]])
    s = (c * dt / dx) * np.sqrt(np.sum(np.sin(k * dx / 2) ** 2))
    w = 2 * np.arcsin(s) / dt
    z, y, x = np.meshgrid(*[np.arange(m, dtype=np.float64) for m in n[::-1]], indexing="ij")
    ph = k[0] * x + k[1] * y + k[2] * z
    f0 = np.cos(ph)
    g0 = (np.cos(ph - w * dt) - np.cos(ph)) / dt
    m = variants.machine("wave3d_f64", n)
    params(m, c, dx, dt)
    m.set("fieldF", f0)
    m.set("fieldG", g0)
    m.run("proceed", steps)
    assert np.abs(m.get("fieldF") - np.cos(ph - w * dt * steps)).max() < 1e-11


@pytest.mark.parametrize("variant", ["wave3d_f64", "wave3d_f64_march"])
def test_wave3d_adjoint_dot_product(require_gpu, variant):
    n = (16, 12, 10)
    nsteps, every = 7, 3
    rng = np.random.default_rng(5)
    x = [torch.tensor(rng.standard_normal(n[::-1]), device="cuda", requires_grad=True)
         for _ in range(2)]
    y = [torch.tensor(rng.standard_normal(n[::-1]), device="cuda") for _ in range(2)]
    m = variants.machine(variant, n)
    params(m)
    out = rollout(m, ["fieldF", "fieldG"], x, nsteps, every)
    lhs = sum((o * yy).sum() for o, yy in zip(out, y))
    lhs.backward()
    rhs = sum((xx.detach() * xx.grad).sum() for xx in x)        # <x, A^T y> with A linear
    assert abs(float(lhs) - float(rhs)) <= 1e-12 * abs(float(lhs))


@pytest.mark.parametrize("variant", ["wave3d_f64", "wave3d_f64_march"])
def test_wave3d_two_step_adjoint_equals_two_adjoint_steps(require_gpu, variant):
    """The rollout driven by the fused two-step kernels (proceed_x2, proceed_adjoint_x2) gives
    the outputs and the gradient of twice as many single steps, bit for bit, and its adjoint
    is the transpose of its forward."""
    n = (40, 9, 70)
    launches, every = 6, 2
    print("wave two-step adjoint %s %s" % (variant, n), flush=True)
    rng = np.random.default_rng(6)
    x0 = [rng.standard_normal(n[::-1]) for _ in range(2)]
    y = [torch.tensor(rng.standard_normal(n[::-1]), device="cuda") for _ in range(2)]
    res = []
    for fwd, adj, k in (("proceed", "proceed_adjoint", 1), ("proceed_x2", "proceed_adjoint_x2", 2)):
        x = [torch.tensor(a, device="cuda", requires_grad=True) for a in x0]
        m = variants.machine(variant, n)
        params(m)
        out = rollout(m, ["fieldF", "fieldG"], x, launches * 2 // k, every * 2 // k, fwd, adj)
        lhs = sum((o * yy).sum() for o, yy in zip(out, y))
        lhs.backward()
        rhs = sum((xx.detach() * xx.grad).sum() for xx in x)
        assert abs(float(lhs) - float(rhs)) <= 1e-12 * abs(float(lhs))
        res.append([o.detach().cpu().numpy() for o in out] + [xx.grad.cpu().numpy() for xx in x])
    for a, b in zip(*res):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("variant,k", [("wave3d_f64", 2), ("wave3d_f64_march", 2),
                                       ("wave3d_f64_persist", 2)])
def test_wave3d_several_steps_per_launch(require_gpu, variant, k):
    """proceed_x<k> (k leapfrog steps per launch, halo k) stores what k x proceed stores."""
    n = (40, 9, 70)
    print("wave x%d %s %s" % (k, variant, n), flush=True)
    rng = np.random.default_rng(8)
    f0, g0 = rng.standard_normal(n[::-1]), rng.standard_normal(n[::-1])
    gpu = variants.machine(variant, n)
    ora = oracle_machine(lambda: add_adjoint(linear_wave_om(3, CType.Double), "proceed",
                                             ["fieldF", "fieldG"]), n)
    for m in (gpu, ora):
        params(m)
        m.set("fieldF", f0)
        m.set("fieldG", g0)
    gpu.run("proceed_x%d" % k, 20 // k)
    ora.run("proceed", 20)
    assert np.array_equal(gpu.get("fieldF"), ora.get("fieldF"))
    assert np.array_equal(gpu.get("fieldG"), ora.get("fieldG"))
    assert abs(gpu.get("energy") - ora.get("energy")) <= 1e-12 * abs(ora.get("energy"))


@pytest.mark.parametrize("k", [2])
def test_wave3d_f32_several_steps_per_launch_within_rounding(require_gpu, k):
    """The FP32 build (contracted multiply-adds) at its tuned streaming tile: k steps per
    launch against k x proceed on the same machine, equal to FP32 rounding (the two forms
    contract different products), 20 steps of a random state."""
    n = (128, 64, 96)
    print("wave f32 x%d %s" % (k, n), flush=True)
    rng = np.random.default_rng(9)
    f0, g0 = rng.standard_normal(n[::-1]), rng.standard_normal(n[::-1])
    a, b = variants.machine("wave3d_f32", n), variants.machine("wave3d_f32", n)
    for m in (a, b):
        params(m)
        m.set("fieldF", f0)
        m.set("fieldG", g0)
    a.run("proceed", 20)
    b.run("proceed_x%d" % k, 20 // k)
    for f in ("fieldF", "fieldG"):
        x, y = a.get(f).astype(np.float64), b.get(f).astype(np.float64)
        assert np.abs(x - y).max() <= 1e-5 * np.abs(x).max(), f


def test_wave3d_f32_two_step_adjoint_within_rounding(require_gpu):
    """The FP32 machine at its tuned streaming tile: the gradient through the two-step pair
    (what the benchmark runs) equals the gradient through single steps to FP32 rounding."""
    n = (128, 64, 96)
    print("wave f32 two-step adjoint %s" % (n,), flush=True)
    rng = np.random.default_rng(10)
    x0 = [rng.standard_normal(n[::-1]).astype(np.float32) for _ in range(2)]
    y = [torch.tensor(rng.standard_normal(n[::-1]).astype(np.float32), device="cuda")
         for _ in range(2)]
    grads = []
    for fwd, adj, k in (("proceed", "proceed_adjoint", 1), ("proceed_x2", "proceed_adjoint_x2", 2)):
        x = [torch.tensor(a, device="cuda", requires_grad=True) for a in x0]
        m = variants.machine("wave3d_f32", n)
        params(m)
        out = rollout(m, ["fieldF", "fieldG"], x, 8 // k, 4 // k, fwd, adj)
        sum((o * yy).sum() for o, yy in zip(out, y)).backward()
        grads.append([xx.grad.double().cpu().numpy() for xx in x])
    for a, b in zip(*grads):
        assert np.abs(a - b).max() <= 1e-5 * np.abs(a).max()


# ---------------------------------------------------------------- the FP32 product machine
WN, WSTEPS = (64, 64, 64), 300


@pytest.fixture(scope="module")
def pulse_oracles():
    """The FP64 and the correctly rounded FP32 oracle after 300 steps of the benchmark's input
    (Gaussian pulse of width 8 cells from seed 3, c = 1, dx = 1, dt = 0.5) on 64^3 cells."""
    from paraiso_amd import inputs
    out = {}
    for real, dt in ((CType.Double, np.float64), (CType.Float, np.float32)):
        print("wave oracle %s %s x %d" % (real.value, WN, WSTEPS), flush=True)
        ora = oracle_machine(lambda: add_adjoint(linear_wave_om(3, real), "proceed",
                                                 ["fieldF", "fieldG"]), WN)
        params(ora)
        ora.set("fieldF", inputs.gaussian_pulse(WN, 3, dtype=dt))
        ora.set("fieldG", 0.0)
        ora.run("proceed", WSTEPS)
        out[real] = (ora.get("fieldF"), ora.get("fieldG"), ora.get("energy"))
        ora.close()
    return out


@pytest.mark.parametrize("kernel,k", [("proceed", 1), ("proceed_x2", 2)])
def test_wave3d_f32_vs_f64_oracle(require_gpu, pulse_oracles, kernel, k):
    """wave3d_f32 (the benchmark's machine, both launch forms) against the FP64 oracle after
    300 steps: fieldF within 5e-6 relative L1 and 2e-6 absolute, fieldG (a velocity near zero
    in most cells) within 5e-5 relative L1, the energy within 1e-6, and each no worse than
    four times the correctly rounded FP32 oracle's own error."""
    from helpers import rel_l1
    from paraiso_amd import inputs
    print("wave f32 %s vs oracle %s" % (kernel, WN), flush=True)
    m = variants.machine("wave3d_f32", WN)
    params(m)
    m.set("fieldF", inputs.gaussian_pulse(WN, 3, dtype=np.float32))
    m.set("fieldG", 0.0)
    m.run(kernel, WSTEPS // k)
    f64, f32 = pulse_oracles[CType.Double], pulse_oracles[CType.Float]
    gF, gG, gE = m.get("fieldF"), m.get("fieldG"), m.get("energy")
    eF, eG = rel_l1(gF, f64[0]), rel_l1(gG, f64[1])
    rF, rG = rel_l1(f32[0], f64[0]), rel_l1(f32[1], f64[1])
    print("rel L1 F %.3g (FP32 oracle %.3g), G %.3g (%.3g)" % (eF, rF, eG, rG))
    assert eF <= 5e-6 and eF <= 4 * rF
    assert eG <= 5e-5 and eG <= 4 * rG
    assert np.abs(gF.astype(np.float64) - f64[0]).max() <= 2e-6
    assert abs(gE - f64[2]) <= 1e-6 * abs(f64[2])


def test_wave3d_f32_adjoint_central_differences(require_gpu):
    """The FP32 gradient of the misfit J = 1/2 sum (f_N - d)^2 (300 steps through the two-step
    kernels and their adjoint, a checkpoint every 20 steps) against central differences of
    the FP32 forward along four seeded directions.  J is quadratic in the initial state, so
    central differences are exact for any step: what remains is FP32 rounding."""
    from paraiso_amd import inputs
    n, steps, every = WN, WSTEPS, 20
    print("wave f32 adjoint FD %s x %d" % (n, steps), flush=True)
    f0 = inputs.gaussian_pulse(n, 3, dtype=np.float32)
    d = inputs.gaussian_pulse(n, 4, dtype=np.float64)
    m = variants.machine("wave3d_f32", n)
    params(m)
    x = [torch.tensor(f0, device="cuda", requires_grad=True),
         torch.zeros(n[::-1], dtype=torch.float32, device="cuda", requires_grad=True)]
    out = rollout(m, ["fieldF", "fieldG"], x, steps // 2, every // 2, "proceed_x2",
                  "proceed_adjoint_x2")
    dd = torch.tensor(d, device="cuda")
    J = 0.5 * ((out[0].double() - dd) ** 2).sum()
    J.backward()
    gF, gG = x[0].grad.double().cpu().numpy(), x[1].grad.double().cpu().numpy()

    def misfit(F0, G0):
        m.set("fieldF", F0)
        m.set("fieldG", G0)
        m.run("proceed_x2", steps // 2)
        r = m.get("fieldF").astype(np.float64) - d
        return 0.5 * float((r * r).sum())
    rng = np.random.default_rng(17)
    for j in range(4):
        dF = (0.2 * rng.standard_normal(n[::-1])).astype(np.float32)
        dG = (0.05 * rng.standard_normal(n[::-1])).astype(np.float32)
        eps = 0.5
        jp = misfit(f0 + eps * dF, eps * dG)
        jm = misfit(f0 - eps * dF, -eps * dG)
        fd = (jp - jm) / (2 * eps)
        ad = float((gF * dF).sum() + (gG * dG).sum())
        print("direction %d: adjoint %.9g, central difference %.9g" % (j, ad, fd))
        assert abs(ad - fd) <= 1e-4 * abs(fd)
