"""The HIP backend's lowering (paraiso_amd/generator/hip/ir.py) on the CPU.

The algebraic identities it applies with one constant operand are checked by
brute force in IEEE arithmetic: every rewrite a correctly rounded build applies
must give the same bits as the operation for every operand (signed zeros,
infinities, NaN, subnormals, random values), in float32 and float64.  The
rewrites kept for fast builds are the ones that fail that check only in the
sign of a zero or for non-finite operands.
"""
import numpy as np
import pytest

from paraiso_amd import annotation as A
from paraiso_amd.generator.hip.ir import _identity, lower_kernel
from paraiso_amd.om import builder as B
from paraiso_amd.om.arith import Op
from paraiso_amd.om.dyn import CType, Realm
from paraiso_amd.optimization import optimize

OPS = {Op.Add: np.add, Op.Sub: np.subtract, Op.Mul: np.multiply, Op.Div: np.divide}


def operands(dt):
    fi = np.finfo(dt)
    special = [0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, np.nan, fi.tiny, -fi.tiny,
               fi.tiny / 4, fi.max, -fi.max, fi.eps]
    rng = np.random.default_rng(7)
    rand = rng.standard_normal(64) * 10.0 ** rng.integers(-30, 30, 64)
    return np.array(special + list(rand), dtype=dt)


def same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    nan = np.isnan(a) & np.isnan(b)
    return bool(np.all(nan | (a.view(np.uint8).reshape(a.size, -1) ==
                              b.view(np.uint8).reshape(b.size, -1)).all(axis=1)))


@pytest.mark.parametrize("dt,ct", [(np.float32, CType.Float), (np.float64, CType.Double)])
@pytest.mark.parametrize("op", list(OPS))
@pytest.mark.parametrize("const", [0.0, -0.0, 1.0, -1.0, 2.0])
@pytest.mark.parametrize("side", [0, 1])
def test_exact_identities_preserve_bits(dt, ct, op, const, side):
    x = operands(dt)
    c = dt(const)
    A_, B_ = (0, 1) if side == 1 else (1, 0)        # value id 0 is x, 1 is the constant
    consts = {1: c}
    r = _identity(op, (A_, B_), consts, ct, fast=False)
    with np.errstate(all="ignore"):
        full = OPS[op](x, c) if side == 1 else OPS[op](c, x)
    if r is None:
        return
    kind, val = r
    assert kind == "arg" and val == 0, r
    assert same_bits(full, x), (op, const, side)


@pytest.mark.parametrize("dt,ct", [(np.float32, CType.Float), (np.float64, CType.Double)])
def test_fast_identities_differ_only_in_zero_sign_or_nonfinite(dt, ct):
    x = operands(dt)
    finite = np.isfinite(x)
    for op in OPS:
        for const in (0.0, 1.0):
            for side in (0, 1):
                pair = (0, 1) if side == 1 else (1, 0)
                r = _identity(op, pair, {1: dt(const)}, ct, fast=True)
                if r is None or r == _identity(op, pair, {1: dt(const)}, ct, fast=False):
                    continue
                with np.errstate(all="ignore"):
                    full = OPS[op](x, dt(const)) if side == 1 else OPS[op](dt(const), x)
                want = x if r[0] == "arg" else np.full_like(x, r[1])
                assert np.array_equal(full[finite], want[finite]), (op, const, side)


def test_integer_identities():
    for op, const, want in ((Op.Add, 0, ("arg", 0)), (Op.Mul, 1, ("arg", 0)),
                            (Op.Mul, 0, ("const", 0)), (Op.Sub, 0, ("arg", 0))):
        assert _identity(op, (0, 1), {1: np.int32(const)}, CType.Int, fast=False) == want


S = B.static("x", Realm.Array, CType.Float)


def _lowered(fast):
    def body():
        x = B.bind(B.load(S))
        z = B.bind(B.imm(0.0, CType.Float))
        y = B.bind((x + z) * 1.0 - z)
        w = B.bind(x * z)
        B.store(S, B.select(x.lt(y), y + w, y + w) / 1.0)
    om = optimize(A.Level.O3, B.make_om("K", [], [S], [("k", body)], ndim=1))
    return lower_kernel(om, 0, (True,), fast=fast)


def test_precise_lowering_keeps_zero_additions():
    k = _lowered(False)
    ops = sorted(n.payload[0].name for n in k.nodes if n.op == "arith")
    # x + 0 and x * 0 stay (they differ from x and 0 for -0, inf and NaN); * 1, - 0,
    # / 1 and the select of equal branches go
    assert ops == ["Add", "Add", "Mul"], ops


def test_fast_lowering_drops_zero_additions_and_weights():
    k = _lowered(True)
    assert [n.op for n in k.nodes if n.op == "arith"] == []
    assert k.fast
    dk = lower_kernel(optimize(A.Level.O3, B.make_om(
        "K", [], [S], [("k", lambda: B.store(S, B.sqrt(B.load(S)) / B.load(S)))], ndim=1)),
        0, (True,), fast=False)
    w = {n.payload[0]: dk.weight(n.id) for n in dk.nodes if n.op == "arith"}
    assert w[Op.Sqrt] > 1 and w[Op.Div] > 1
