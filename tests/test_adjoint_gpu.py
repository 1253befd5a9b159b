"""Checkpointed adjoint of fixed-dt Hydro on the GPU (through the C ABI and the
torch.autograd wrapper) against the CPU FP64 reference adjoint (torch autograd
through oracle/hydro_torch.py, an independent restatement of HydroMain.hs) and
central finite differences of the GPU forward."""
import numpy as np
import pytest
import torch

from oracle import hydro_torch as HT
from paraiso_amd import inputs, variants
from paraiso_amd.autograd import rollout

pytestmark = pytest.mark.gpu
F = ("density", "velocity0", "velocity1", "pressure")


def generic_state(nx, ny, seed, amp=0.05):
    rng = np.random.default_rng(seed)
    return [a + amp * rng.standard_normal(a.shape) for a in inputs.kh_init(nx, ny, seed)]


def machine(variant, nx, ny, dt):
    m = variants.machine(variant, (nx, ny))
    for k, v in inputs.hydro_scalars(nx, ny).items():
        m.set(k, v)
    m.set("dt", dt)
    return m


def objective(fields, mask):
    rho, v0, v1, _ = fields
    return (0.5 * rho * (v0 * v0 + v1 * v1) * mask).sum()


def cpu_reference_gradient(state, nx, ny, nsteps, dt):
    t = [torch.tensor(a, dtype=torch.float64, requires_grad=True) for a in state]
    out, _ = HT.run(t, nsteps, (1.0 / nx, 1.0 / ny), 1.4, dt=dt)
    J, _ = HT.kinetic_energy_in_square(*out[:3])
    J.backward()
    return float(J), [x.grad.numpy() for x in t], out


@pytest.mark.parametrize("nsteps,every", [(10, 4), (6, 6)])
def test_gpu_adjoint_matches_cpu_fp64_reference(require_gpu, nsteps, every):
    nx, ny, dt = 32, 32, 1e-3
    print("adjoint case nsteps=%d every=%d" % (nsteps, every), flush=True)
    state = generic_state(nx, ny, 2)
    J_ref, g_ref, out_ref = cpu_reference_gradient(state, nx, ny, nsteps, dt)
    m = machine("hydro_adj_f64", nx, ny, dt)
    fields = [torch.tensor(a, dtype=torch.float64, device="cuda", requires_grad=True) for a in state]
    out = rollout(m, F, fields, nsteps, every)
    _, mask = HT.kinetic_energy_in_square(*[o.detach().cpu() for o in out[:3]])
    J = objective(out, mask.cuda().to(torch.float64))
    J.backward()
    assert abs(float(J) - J_ref) <= 1e-12 * abs(J_ref)
    for f, g, gr in zip(F, fields, g_ref):
        err = np.abs(g.grad.cpu().numpy() - gr).max() / np.abs(gr).max()
        assert err < 1e-10, (f, err)


def test_gpu_adjoint_central_differences(require_gpu):
    nx, ny, dt, nsteps, every = 24, 24, 1e-3, 5, 2
    state = generic_state(nx, ny, 5)
    m = machine("hydro_adj_f64", nx, ny, dt)
    x = [torch.tensor(a, dtype=torch.float64, device="cuda", requires_grad=True) for a in state]
    mask = HT.kinetic_energy_in_square(torch.zeros(ny, nx, dtype=torch.float64),
                                       torch.zeros(ny, nx, dtype=torch.float64),
                                       torch.zeros(ny, nx, dtype=torch.float64))[1]
    mask = mask.to(torch.float64).cuda()
    J = objective(rollout(m, F, x, nsteps, every), mask)
    J.backward()
    grads = [xi.grad.detach() for xi in x]
    rng = np.random.default_rng(16)
    eps = 1e-7          # limiter kinks (minmod, max floors, HLLC selects) spoil larger steps
    worst = 0.0
    for k in range(16):
        d = [torch.tensor(rng.standard_normal((ny, nx)), device="cuda") for _ in F]
        with torch.no_grad():
            jp = objective(rollout(m, F, [a + eps * b for a, b in zip(x, d)], nsteps, every), mask)
            jm = objective(rollout(m, F, [a - eps * b for a, b in zip(x, d)], nsteps, every), mask)
        fd = float((jp - jm) / (2 * eps))
        an = float(sum((g * b).sum() for g, b in zip(grads, d)))
        worst = max(worst, abs(fd - an) / max(abs(an), 1e-3))
    print("central differences: worst relative gap %.3g over 16 directions" % worst)
    assert worst < 1e-5


@pytest.mark.parametrize("variant", ["hydro_adj_f32", "hydro_adj_f32_fast"])
def test_fp32_adjoint_close_to_fp64(require_gpu, variant):
    nx, ny, dt, nsteps, every = 32, 32, 1e-3, 8, 4
    state = generic_state(nx, ny, 9)
    _, g_ref, _ = cpu_reference_gradient(state, nx, ny, nsteps, dt)
    m = machine(variant, nx, ny, dt)
    fields = [torch.tensor(a, dtype=torch.float32, device="cuda", requires_grad=True) for a in state]
    out = rollout(m, F, fields, nsteps, every)
    _, mask = HT.kinetic_energy_in_square(*[o.detach().cpu().double() for o in out[:3]])
    objective(out, mask.cuda().to(torch.float32)).backward()
    g = fields[0].grad.double().cpu().numpy()
    err = np.abs(g - g_ref[0]).sum() / np.abs(g_ref[0]).sum()
    assert err < 1e-3, err


# ---------------------------------------------------------------- the adjoint per stage
STAGED = dict(fwd=["proceed_stage1", "proceed_stage2"],
              adj=["proceed_stage2_adjoint", "proceed_stage1_adjoint"], tape_after=1)


def test_stages_compose_to_proceed(require_gpu):
    """proceed_stage1 then proceed_stage2 store what proceed stores, bit for bit (correctly
    rounded FP64: the same operations on the same values)."""
    nx, ny, dt, steps = 96, 64, 1e-3, 7
    print("stages vs proceed %dx%d" % (nx, ny), flush=True)
    state = generic_state(nx, ny, 3)
    a, b = machine("hydro_adj2_f64", nx, ny, dt), machine("hydro_adj2_f64", nx, ny, dt)
    for m in (a, b):
        for f, arr in zip(F, state):
            m.set(f, arr)
    a.run("proceed", steps)
    for _ in range(steps):
        b.run("proceed_stage1")
        b.run("proceed_stage2")
    assert a.get("time") == b.get("time")
    for f in F:
        assert np.array_equal(a.get(f), b.get(f)), f


@pytest.mark.parametrize("nsteps,every", [(10, 4), (6, 6), (7, 1)])
def test_staged_adjoint_matches_cpu_fp64_reference(require_gpu, nsteps, every):
    """The adjoint per stage, driven by the tape-rotating checkpoint replay, against the CPU
    FP64 reference adjoint."""
    nx, ny, dt = 32, 32, 1e-3
    print("staged adjoint nsteps=%d every=%d" % (nsteps, every), flush=True)
    state = generic_state(nx, ny, 2)
    J_ref, g_ref, _ = cpu_reference_gradient(state, nx, ny, nsteps, dt)
    m = machine("hydro_adj2_f64", nx, ny, dt)
    fields = [torch.tensor(a, dtype=torch.float64, device="cuda", requires_grad=True)
              for a in state]
    out = rollout(m, F, fields, nsteps, every, **STAGED)
    _, mask = HT.kinetic_energy_in_square(*[o.detach().cpu() for o in out[:3]])
    J = objective(out, mask.cuda().to(torch.float64))
    J.backward()
    assert abs(float(J) - J_ref) <= 1e-12 * abs(J_ref)
    for f, g, gr in zip(F, fields, g_ref):
        err = np.abs(g.grad.cpu().numpy() - gr).max() / np.abs(gr).max()
        assert err < 1e-10, (f, err)


def test_staged_fp32_adjoint_close_to_fp64(require_gpu):
    nx, ny, dt, nsteps, every = 32, 32, 1e-3, 8, 4
    state = generic_state(nx, ny, 9)
    _, g_ref, _ = cpu_reference_gradient(state, nx, ny, nsteps, dt)
    m = machine("hydro_adj2_f32_fast", nx, ny, dt)
    fields = [torch.tensor(a, dtype=torch.float32, device="cuda", requires_grad=True)
              for a in state]
    out = rollout(m, F, fields, nsteps, every, **STAGED)
    _, mask = HT.kinetic_energy_in_square(*[o.detach().cpu().double() for o in out[:3]])
    objective(out, mask.cuda().to(torch.float32)).backward()
    for f, x, gr in zip(F, fields, g_ref):
        g = x.grad.double().cpu().numpy()
        err = np.abs(g - gr).sum() / np.abs(gr).sum()
        assert err < 1e-3, (f, err)


def test_rollouts_of_one_machine_do_not_mix(require_gpu):
    """A second rollout of the same machine overwrites the first one's checkpoints: the
    first one's backward raises instead of returning a gradient of the other run."""
    nx, ny, dt = 24, 24, 1e-3
    m = machine("hydro_adj_f64", nx, ny, dt)
    x1 = [torch.tensor(a, device="cuda", requires_grad=True) for a in generic_state(nx, ny, 1)]
    x2 = [torch.tensor(a, device="cuda", requires_grad=True) for a in generic_state(nx, ny, 2)]
    o1 = rollout(m, F, x1, 3, 2)
    o2 = rollout(m, F, x2, 3, 2)
    o2[0].sum().backward()
    assert x2[0].grad is not None
    with pytest.raises(RuntimeError, match="overwritten"):
        o1[0].sum().backward()


def test_rollout_with_strided_grad_and_transposed_input(require_gpu):
    """J = out[0].sum() gives a stride-0 (expanded) gradient, and a transposed input is
    not contiguous: both are made contiguous on torch's stream, which the machine shares."""
    nx, ny, dt, nsteps, every = 24, 24, 1e-3, 4, 2
    state = generic_state(nx, ny, 4)
    m = machine("hydro_adj_f64", nx, ny, dt)
    base = [torch.tensor(a.T.copy(), device="cuda") for a in state]     # (nx, ny) storage
    x = [b.t().requires_grad_() for b in base]                           # (ny, nx) views
    out = rollout(m, F, x, nsteps, every)
    out[0].sum().backward()
    # the same J through the CPU FP64 reference
    t = [torch.tensor(a, dtype=torch.float64, requires_grad=True) for a in state]
    o, _ = HT.run(t, nsteps, (1.0 / nx, 1.0 / ny), 1.4, dt=dt)
    o[0].sum().backward()
    for f, xi, ti in zip(F, x, t):
        err = np.abs(xi.grad.cpu().numpy() - ti.grad.numpy()).max() / max(
            np.abs(ti.grad.numpy()).max(), 1e-300)
        assert err < 1e-10, (f, err)


def test_joined_adjoint_is_bit_identical(require_gpu):
    """The adjoint kernels with their shared expansions joined into earlier phases compute
    the same correctly rounded operations on the same values: the gradient is bit for bit
    that of the schedule without the join."""
    nx, ny, dt, nsteps, every = 48, 40, 1e-3, 6, 3
    print("joined staged adjoint %dx%d" % (nx, ny), flush=True)
    state = generic_state(nx, ny, 5)
    grads = []
    for variant in ("hydro_adj2_f64", "hydro_adj2_f64_join"):
        m = machine(variant, nx, ny, dt)
        fields = [torch.tensor(a, dtype=torch.float64, device="cuda", requires_grad=True)
                  for a in state]
        out = rollout(m, F, fields, nsteps, every, **STAGED)
        _, mask = HT.kinetic_energy_in_square(*[o.detach().cpu() for o in out[:3]])
        objective(out, mask.cuda().to(torch.float64)).backward()
        grads.append([x.grad.cpu().numpy() for x in fields])
    for f, a, b in zip(F, *grads):
        assert np.array_equal(a, b), f


def test_staged_adjoint_at_the_benchmark_size(require_gpu):
    """The benchmark's adjoint: 512^2 KH state (seed 7), fixed dt 1e-4, 200 steps, a
    checkpoint every 20, J = kinetic energy over [0.25, 0.75]^2 at the final time.

    The FP32 product machine's density gradient against the FP64 machine's (relative L1
    3.4e-5 measured, bound 1e-4); then the FP64
    gradient against central differences of the FP64 forward along 16 seeded random
    directions of the initial density.  200 steps over 512^2 cells cross limiter and HLLC
    switches (and the KH state's exact jumps put the minmod limiter on its kink): where a
    perturbed run crosses one, central differences average the two one-sided derivatives
    while the adjoint takes one branch.  The share of directions that cross one falls with
    the step: over these 16 directions the worst gap was 1.2e-3 at a step of 1e-7 and
    2.8e-4 at 1e-8, the step used here (the median of the first eight: 2.2e-5 and 4.5e-6)."""
    n, dt, nsteps, every = 512, 1e-4, 200, 20
    print("staged adjoint %d^2 x %d" % (n, nsteps), flush=True)
    state = inputs.kh_init(n, n, 7)
    c = (torch.arange(n, dtype=torch.float64) + 0.5) / n
    inside = ((c >= 0.25) & (c <= 0.75)).to(torch.float64)
    mask = (inside[:, None] * inside[None, :]).cuda()
    m64 = machine("hydro_adj2_f64", n, n, dt)
    x = [torch.tensor(a, dtype=torch.float64, device="cuda", requires_grad=True) for a in state]
    objective(rollout(m64, F, x, nsteps, every, **STAGED), mask).backward()
    g64 = x[0].grad.detach()
    m32 = machine("hydro_adj2_f32_fast", n, n, dt)
    x32 = [torch.tensor(a, dtype=torch.float32, device="cuda", requires_grad=True) for a in state]
    objective(rollout(m32, F, x32, nsteps, every, **STAGED), mask.float()).backward()
    g32 = x32[0].grad.double()
    err = float((g32 - g64).abs().sum() / g64.abs().sum())
    print("FP32 density gradient: relative L1 %.3g against FP64" % err, flush=True)
    rng = np.random.default_rng(17)
    eps, worst = 1e-8, 0.0
    base = [xi.detach() for xi in x]
    for k in range(16):
        d = torch.tensor(rng.standard_normal((n, n)), device="cuda")
        with torch.no_grad():
            jp = objective(rollout(m64, F, [base[0] + eps * d] + base[1:], nsteps, every,
                                   **STAGED), mask)
            jm = objective(rollout(m64, F, [base[0] - eps * d] + base[1:], nsteps, every,
                                   **STAGED), mask)
        fd = float((jp - jm) / (2 * eps))
        an = float((g64 * d).sum())
        worst = max(worst, abs(fd - an) / abs(an))
    print("central differences at %d^2: worst relative gap %.3g" % (n, worst), flush=True)
    assert err < 1e-4, err
    assert worst < 5e-4, worst


def test_staged_adjoint_full_grid_matches_cpu_reference(require_gpu):
    """The staged FP64 adjoint on the benchmark's 512^2 grid (a noisy KH state, so that no
    limiter sits exactly on its kink; 4 steps of dt 1e-4, a checkpoint every 2) against the
    CPU FP64 reference adjoint (torch autograd through oracle/hydro_torch.py)."""
    n, dt, nsteps, every = 512, 1e-4, 4, 2
    print("staged adjoint %d^2 x %d vs CPU reference" % (n, nsteps), flush=True)
    state = generic_state(n, n, 7)
    J_ref, g_ref, _ = cpu_reference_gradient(state, n, n, nsteps, dt)
    m = machine("hydro_adj2_f64", n, n, dt)
    x = [torch.tensor(a, dtype=torch.float64, device="cuda", requires_grad=True) for a in state]
    out = rollout(m, F, x, nsteps, every, **STAGED)
    _, mask = HT.kinetic_energy_in_square(*[o.detach().cpu() for o in out[:3]])
    J = objective(out, mask.cuda().to(torch.float64))
    J.backward()
    assert abs(float(J) - J_ref) <= 1e-12 * abs(J_ref)
    for f, g, gr in zip(F, x, g_ref):
        err = np.abs(g.grad.cpu().numpy() - gr).max() / np.abs(gr).max()
        assert err < 1e-10, (f, err)
