"""Products by folded constants in FP32 builds (paraiso_amd/generator/hip/emit.py).

A product x * c by a constant that is no binary32 value is emitted compensated,
fma(x, hi, x * lo), when the kernel also multiplies by 1/c: the Hydro step converts the
pressure to the energy and back through gamma - 1 (Hydro.hs:99-127), and a product by
the rounded 0.4f would make that round trip drift by the same factor every step.  Other
constants (gamma in the sound speed, (gamma + 1) / (2 gamma) in the wave-speed estimate,
HydroMain.hs:237-276) are multiplied by their nearest binary32 value."""
from fractions import Fraction

import numpy as np

from paraiso_amd import variants
from paraiso_amd.generator.hip.emit import _const_product, _split_const, inverse_pairs
from paraiso_amd.generator.hip.ir import lower_kernel
from paraiso_amd.generator.plan import translate


def _kernel(variant, name):
    mk, _, precise = variants.VARIANTS[variant]
    om = translate(variants.setup_of(variant), mk()).om
    ki = [k.name for k in om.kernels].index(name)
    return lower_kernel(om, ki, (True,) * om.ndim, exact=True)


def test_hydro_pairs_are_gamma_minus_one_and_its_reciprocal():
    assert inverse_pairs(_kernel("hydro_f32", "proceed_pipelined")) == {Fraction(2, 5),
                                                                         Fraction(5, 2)}


def test_wave_pairs_are_exact():
    """LinearWave multiplies by 2 and by 1/2 (the Laplacian, the mean of g0 and g1): a pair,
    but both are binary32 values, so no product is compensated."""
    pairs = inverse_pairs(_kernel("wave3d_f32", "proceed_x2"))
    assert pairs == {Fraction(1, 2), Fraction(2)}
    assert all(_split_const(c)[1] == 0 for c in pairs)


def test_products_compensated_only_for_the_pairs():
    pairs = {Fraction(2, 5), Fraction(5, 2)}
    assert _const_product("x", Fraction(2, 5), pairs).startswith("__builtin_fmaf(x, 0.4f,")
    assert _const_product("x", Fraction(7, 5), pairs) == "(x) * 1.4f"
    assert _const_product("x", Fraction(5, 2), pairs) == "(x) * 2.5f"       # exact anyway
    assert _const_product("x", Fraction(7, 5)).startswith("__builtin_fmaf(")   # all


def test_split_constant_is_exact_to_48_bits():
    for c in (Fraction(2, 5), Fraction(7, 5), Fraction(6, 7)):
        hi, lo = _split_const(c)
        err = abs(Fraction(float(hi)) + Fraction(float(lo)) - c) / c
        assert err < Fraction(1, 2 ** 47)
        assert np.float32(hi) == hi and abs(float(lo)) < abs(float(hi)) * 2 ** -23
