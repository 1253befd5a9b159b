"""Pins of the Hydro oracle (the PlanTrans restatement running the current
HydroMain.hs program) by what the mathematics fixes, on the CPU:

* the exact Riemann solver used as the shock-tube reference is itself checked
  against Toro's published star states and the jump relations of its waves;
* double Sod shock tube on a periodic line: the scheme converges to the exact
  solution, with the shock where the Rankine-Hugoniot speed puts it;
* entropy wave: a density profile advected at constant velocity and pressure
  converges at second order;
* sound wave: a small right-going acoustic wave travels at sqrt(gamma p / rho);
* mass, momentum and energy are conserved to roundoff on periodic grids
  (flux form, HydroMain.hs:185-196) while the floors are inactive;
* the x<->y transposed run mirrors the run;
* an independent torch restatement written from the equations agrees.
A dropped flux term, a wrong sign, a swapped axis or a wrong sound speed each
fails at least one of these.
"""
from fractions import Fraction
import math

import numpy as np
import pytest
import torch

from helpers import oracle_machine
from oracle import hydro_torch as HT
from oracle import riemann_exact as RE
from paraiso_amd import inputs
from paraiso_amd.om.dyn import CType
from paraiso_amd.schemes.hydro import hydro_om

F = ("density", "velocity0", "velocity1", "pressure")
G = 1.4
SOD_L = RE.Side(1.0, 0.0, 1.0)
SOD_R = RE.Side(0.125, 0.0, 0.1)


def machine(nx, ny):
    return oracle_machine(lambda: hydro_om(CType.Double, Fraction(7, 5)), (nx, ny))


def load(m, nx, ny, state, cfl=0.4):
    for k, v in inputs.hydro_scalars(nx, ny, cfl=cfl).items():
        m.set(k, v)
    for name, arr in zip(F, state):
        m.set(name, np.ascontiguousarray(arr, dtype=np.float64))


def run_until(m, t_end, max_steps=100000):
    steps = 0
    while m.get("time") < t_end and steps < max_steps:
        m.run("proceed", 1)
        steps += 1
    return m.get("time")


def line_state(rho, u, p, ny):
    """1-D profiles along x repeated over ny rows; velocity1 = 0."""
    rep = lambda a: np.repeat(np.asarray(a, dtype=np.float64)[None, :], ny, axis=0)
    return rep(rho), rep(u), np.zeros((ny, len(rho))), rep(p)


# --------------------------------------------------------------------- the exact solver
@pytest.mark.parametrize("L,R,ref", [
    # Toro, Riemann Solvers, table 4.3: p*, u*, rho*L, rho*R (tests 1, 3, 4, 5)
    ((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), (0.30313, 0.92745, 0.42632, 0.26557)),
    ((1.0, 0.0, 1000.0), (1.0, 0.0, 0.01), (460.894, 19.5975, 0.57506, 5.99924)),
    ((1.0, 0.0, 0.01), (1.0, 0.0, 100.0), (46.0950, -6.19633, 5.99242, 0.57511)),
    ((5.99924, 19.5975, 460.894), (5.99242, -6.19633, 46.0950),
     (1691.64, 8.68975, 14.2823, 31.0426)),
])
def test_exact_solver_star_states_match_toro(L, R, ref):
    Ls, Rs = RE.Side(*L), RE.Side(*R)
    p, u = RE.star(Ls, Rs, G)
    rl, rr = RE.star_densities(Ls, Rs, G, p)
    for got, want in zip((p, u, rl, rr), ref):
        assert abs(got - want) <= 6e-5 * max(1.0, abs(want)), (got, want)


def test_exact_solver_waves_satisfy_jump_relations():
    p, u = RE.star(SOD_L, SOD_R, G)
    rl, rr = RE.star_densities(SOD_L, SOD_R, G, p)
    # right shock: mass and momentum flux continuous in the shock frame
    S = RE.shock_speed(SOD_R, G, p, right=True)
    assert abs(SOD_R.rho * (SOD_R.u - S) - rr * (u - S)) < 1e-12
    assert abs(SOD_R.rho * SOD_R.u * (SOD_R.u - S) + SOD_R.p - (rr * u * (u - S) + p)) < 1e-12
    eR = SOD_R.p / (G - 1) + 0.5 * SOD_R.rho * SOD_R.u ** 2
    eS = p / (G - 1) + 0.5 * rr * u * u
    assert abs((eR + SOD_R.p) * SOD_R.u - eR * S - ((eS + p) * u - eS * S)) < 1e-12
    # left rarefaction: isentropic, and the Riemann invariant u + 2c/(g-1) is constant
    assert abs(SOD_L.p / SOD_L.rho ** G - p / rl ** G) < 1e-12
    cL, cS = math.sqrt(G * SOD_L.p / SOD_L.rho), math.sqrt(G * p / rl)
    assert abs(SOD_L.u + 2 * cL / (G - 1) - (u + 2 * cS / (G - 1))) < 1e-12
    # the sampled fan is continuous at its head and tail
    head, tail = SOD_L.u - cL, u - cS
    r, uu, pp = RE.sample(SOD_L, SOD_R, G, [head - 1e-9, head + 1e-9, tail - 1e-9, tail + 1e-9])
    assert abs(r[0] - r[1]) < 1e-6 and abs(r[2] - r[3]) < 1e-6 and abs(pp[2] - pp[3]) < 1e-6


# --------------------------------------------------------------------- shock tube
def double_sod(nx, t_end, ny=4):
    """High state on [0.25, 0.75), low state elsewhere: a Sod problem at x = 0.75 and its
    mirror image at x = 0.25; their waves do not meet before t = 0.14."""
    x = (np.arange(nx) + 0.5) / nx
    hi = (x >= 0.25) & (x < 0.75)
    rho = np.where(hi, SOD_L.rho, SOD_R.rho)
    p = np.where(hi, SOD_L.p, SOD_R.p)
    m = machine(nx, ny)
    load(m, nx, ny, line_state(rho, np.zeros(nx), p, ny))
    t = run_until(m, t_end)
    # exact: cell averages of the two self-similar solutions (8 samples per cell)
    sub = (np.arange(nx * 8) + 0.5) / (nx * 8)
    right = sub >= 0.5
    xi = np.where(right, (sub - 0.75) / t, (0.25 - sub) / t)
    r_ex, u_ex, p_ex = RE.sample(SOD_L, SOD_R, G, xi)
    u_ex = np.where(right, u_ex, -u_ex)
    avg = lambda a: a.reshape(nx, 8).mean(axis=1)
    return m, t, avg(r_ex), avg(u_ex), avg(p_ex)


def test_double_sod_converges_to_exact_solution():
    errs = {}
    for nx in (200, 400):
        m, t, r_ex, u_ex, p_ex = double_sod(nx, 0.1)
        rho = m.get("density")
        assert np.abs(rho - rho[0:1]).max() == 0.0          # rows stay identical
        errs[nx] = float(np.abs(rho[0] - r_ex).mean())
        # the right-moving shock sits where its Rankine-Hugoniot speed puts it
        p_star, u_star = RE.star(SOD_L, SOD_R, G)
        _, rr = RE.star_densities(SOD_L, SOD_R, G, p_star)
        xs = 0.75 + RE.shock_speed(SOD_R, G, p_star, right=True) * t
        x = (np.arange(nx) + 0.5) / nx
        zone = (x > 0.8) & (x < 0.99)
        mid = 0.5 * (rr + SOD_R.rho)
        cross = x[zone][np.argmax(rho[0][zone] < mid)]
        assert abs(cross - xs) < 3.0 / nx, (nx, cross, xs)
        # the star plateau between contact and shock carries p* and u*
        pl = (x > 0.75 + u_star * t + 4.0 / nx) & (x < xs - 4.0 / nx)
        assert np.abs(m.get("pressure")[0][pl] - p_star).max() < 0.02 * p_star
        assert np.abs(m.get("velocity0")[0][pl] - u_star).max() < 0.02 * u_star
    print("double Sod L1(rho):", errs)
    assert errs[200] < 0.012 and errs[400] < 0.007
    assert errs[400] < 0.75 * errs[200]


# --------------------------------------------------------------------- smooth waves
def test_entropy_wave_second_order():
    errs = []
    for nx in (64, 128):
        x = (np.arange(nx) + 0.5) / nx
        rho = 1.0 + 0.2 * np.sin(2 * np.pi * x)
        m = machine(nx, 4)
        load(m, nx, 4, line_state(rho, np.ones(nx), np.ones(nx), 4))
        t = run_until(m, 0.5)
        # exact: the cell average of rho advected by t
        exact = 1.0 + 0.2 * (np.cos(2 * np.pi * (x - 0.5 / nx - t)) -
                             np.cos(2 * np.pi * (x + 0.5 / nx - t))) / (2 * np.pi / nx)
        errs.append(float(np.abs(m.get("density")[0] - exact).mean()))
        assert np.abs(m.get("velocity0")[0] - 1.0).max() < 1e-12
        assert np.abs(m.get("pressure")[0] - 1.0).max() < 1e-12
    order = math.log2(errs[0] / errs[1])
    print("entropy wave L1:", errs, "order %.2f" % order)
    assert errs[1] < 2e-3 and order > 1.6


def test_sound_wave_travels_at_sound_speed():
    nx, eps = 256, 1e-4
    c0 = math.sqrt(G)
    x = (np.arange(nx) + 0.5) / nx
    d = eps * np.sin(2 * np.pi * x)
    m = machine(nx, 4)
    load(m, nx, 4, line_state(1.0 + d, c0 * d, 1.0 + c0 * c0 * d, 4))
    t = run_until(m, 0.25)
    r = m.get("density")[0] - 1.0
    phase = np.angle(np.sum(r * np.exp(-2j * np.pi * x)))    # profile ~ sin(2 pi (x - ct))
    speed = ((-np.pi / 2 - phase) % (2 * np.pi)) / (2 * np.pi) / t
    print("sound speed measured %.6f expected %.6f" % (speed, c0))
    assert abs(speed - c0) < 2e-3 * c0
    amp = 2 * np.abs(np.sum(r * np.exp(-2j * np.pi * x))) / nx
    assert 0.9 * eps < amp <= 1.0001 * eps


# --------------------------------------------------------------------- invariants
def conserved_totals(m):
    r, u, v, p = (m.get(f) for f in F)
    return np.array([r.sum(), (r * u).sum(), (r * v).sum(),
                     (p / (G - 1) + 0.5 * r * (u * u + v * v)).sum()])


def test_periodic_conservation_to_roundoff():
    nx, ny = 64, 48
    m = machine(nx, ny)
    load(m, nx, ny, inputs.kh_init(nx, ny, 21))
    before = conserved_totals(m)
    m.run("proceed", 20)
    after = conserved_totals(m)
    scale = np.array([before[0], np.abs(m.get("density") * m.get("velocity0")).sum(),
                      np.abs(m.get("density") * m.get("velocity1")).sum(), before[3]])
    rel = np.abs(after - before) / scale
    print("conservation drift:", rel)
    assert (rel < 1e-13).all(), rel
    assert m.get("density").min() > 0.5 and m.get("pressure").min() > 1.0   # floors inactive


def test_transposed_run_mirrors():
    nx, ny = 40, 40
    s = inputs.kh_init(nx, ny, 8)
    a = machine(nx, ny)
    load(a, nx, ny, s)
    b = machine(ny, nx)
    load(b, ny, nx, (s[0].T, s[2].T, s[1].T, s[3].T))
    a.run("proceed", 10)
    b.run("proceed", 10)
    assert abs(a.get("time") - b.get("time")) < 1e-15
    pairs = (("density", "density"), ("velocity0", "velocity1"), ("velocity1", "velocity0"),
             ("pressure", "pressure"))
    for fa, fb in pairs:
        err = np.abs(a.get(fa) - b.get(fb).T).max()
        assert err < 1e-12, (fa, err)


def test_torch_restatement_agrees_with_oracle():
    nx, ny, steps = 48, 32, 8
    s = inputs.kh_init(nx, ny, 4)
    m = machine(nx, ny)
    load(m, nx, ny, s)
    m.run("proceed", steps)
    with torch.no_grad():
        out, t = HT.run([torch.tensor(a) for a in s], steps, (1.0 / nx, 1.0 / ny), G, cfl=0.4)
    assert abs(float(t) - m.get("time")) < 1e-14
    for f, o in zip(F, out):
        err = np.abs(o.numpy() - m.get(f)).max()
        assert err < 1e-13, (f, err)
