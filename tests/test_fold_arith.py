"""The arithmetic of the one-word integer-sum finish (the epilogue the HIP emitter writes for
a launch whose one reduction is an integer sum, paraiso_amd/generator/hip/emit.py): every
workgroup adds (1 << 48) + (partial ^ 0x80000000) to a 64-bit word; the one whose fetched
value counts nwg - 1 arrivals recovers the wrapping int32 sum of all partials.  Checked here
on the CPU, against Python's exact sum mod 2^32, for partials of both signs and the extremes,
in arrival orders of every kind; the GPU runs it in Life's bit-exact population tests."""
import random

import pytest

MASK64 = (1 << 64) - 1


def fold(partials):
    """Simulate the atomics in the given arrival order; returns (sum seen by the last
    workgroup, the word it leaves)."""
    nwg = len(partials)
    word = 0
    result = None
    for p in partials:
        u = (p & 0xFFFFFFFF) ^ 0x80000000
        old = word
        word = (word + (1 << 48) + u) & MASK64
        if (old >> 48) == nwg - 1:
            tot = (old & 0xFFFFFFFFFFFF) + u
            r = (tot - nwg * 0x80000000) & 0xFFFFFFFF
            result = r - (1 << 32) if r >= 1 << 31 else r
            word = 0                      # the last workgroup clears the word
    return result, word


def wrap32(x):
    x &= 0xFFFFFFFF
    return x - (1 << 32) if x >= 1 << 31 else x


@pytest.mark.parametrize("nwg", [1, 2, 3, 512, 4096, 65535])
def test_last_arrival_has_the_wrapping_sum(nwg):
    rng = random.Random(nwg)
    kinds = [
        [rng.randrange(-2 ** 31, 2 ** 31) for _ in range(nwg)],
        [2 ** 31 - 1] * nwg,
        [-2 ** 31] * nwg,
        [rng.choice((-2 ** 31, 2 ** 31 - 1, 0, -1, 1)) for _ in range(nwg)],
        [rng.randrange(0, 2049) for _ in range(nwg)],      # Life: a tile's live cells
    ]
    for parts in kinds:
        got, word = fold(parts)
        assert got == wrap32(sum(parts))
        assert word == 0


def test_partial_field_never_carries_into_the_count():
    """65,535 partials of at most 2^32 - 1 each stay below 2^48."""
    assert 65535 * (2 ** 32 - 1) < 2 ** 48
