"""Seeded synthetic inputs for the benchmark configurations.

All randomness is a counter-based generator (splitmix64 of ``seed ^ index``),
so any cell's value can be reproduced independently of the others, on any
platform.  The same functions feed the GPU machines, the oracle and the
fixtures.
"""
from __future__ import annotations

import math

import numpy as np

_M1 = np.uint64(0x9E3779B97F4A7C15)
_M2 = np.uint64(0xBF58476D1CE4E5B9)
_M3 = np.uint64(0x94D049BB133111EB)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser of ``x + golden`` (Steele, Lea & Flood 2014)."""
    with np.errstate(over="ignore"):
        z = x.astype(np.uint64) + _M1
        z = (z ^ (z >> np.uint64(30))) * _M2
        z = (z ^ (z >> np.uint64(27))) * _M3
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """n doubles in [0, 1) from the top 53 bits of splitmix64(seed ^ (stream<<40 | i))."""
    idx = np.arange(n, dtype=np.uint64) | (np.uint64(stream) << np.uint64(40))
    bits = splitmix64(idx ^ np.uint64(seed & 0xFFFFFFFFFFFFFFFF))
    return (bits >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def life_init(shape, seed: int) -> np.ndarray:
    """Bernoulli(0.5) int32 cells; shape (ny, nx); bit 63 of splitmix64(seed ^ i)."""
    n = int(np.prod(shape))
    bits = splitmix64(np.arange(n, dtype=np.uint64) ^ np.uint64(seed))
    return (bits >> np.uint64(63)).astype(np.int32).reshape(shape)


def kh_init(nx: int, ny: int, seed: int, dtype=np.float64, band=(0.25, 0.75),
            rho_in=2.0, rho_out=1.0, v_in=0.5, v_out=-0.5, p0=2.5, amp=0.01):
    """Kelvin-Helmholtz state on the unit square, cell centres y = (j + 0.5)/ny.

    density = rho_in in band[0] <= y < band[1], else rho_out; velocity0 = v_in
    in the band, else v_out; pressure = p0; velocity1 = amp * U(-1, 1) per cell.
    Returns (density, velocity0, velocity1, pressure), each (ny, nx).
    """
    y = (np.arange(ny, dtype=np.float64) + 0.5) / ny
    inband = (y >= band[0]) & (y < band[1])
    dens = np.where(inband, rho_in, rho_out)[:, None] * np.ones((1, nx))
    v0 = np.where(inband, v_in, v_out)[:, None] * np.ones((1, nx))
    noise = uniform01(seed, nx * ny).reshape(ny, nx) * 2.0 - 1.0
    v1 = amp * noise
    p = np.full((ny, nx), p0)
    return tuple(np.ascontiguousarray(a.astype(dtype)) for a in (dens, v0, v1, p))


def hydro_scalars(nx: int, ny: int, cfl: float = 0.4):
    return {"dR0": 1.0 / nx, "dR1": 1.0 / ny, "extent0": 1.0, "extent1": 1.0,
            "cfl": cfl, "time": 0.0, "generation": 0}


def gaussian_pulse(shape, seed: int, width: float = 8.0, amp: float = 1.0, dtype=np.float32):
    """Periodic Gaussian bump of ``width`` cells centred at a seeded cell; shape (nz, ny, nx)."""
    c = uniform01(seed, len(shape), stream=7) * np.array(shape)
    grids = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")
    r2 = np.zeros(shape)
    for ax, n in enumerate(shape):
        d = grids[ax] - c[ax]
        d = (d + n / 2.0) % n - n / 2.0        # periodic distance
        r2 += d * d
    return (amp * np.exp(-r2 / (2.0 * width * width))).astype(dtype)


def batched_sizes(count: int, seed: int, lo: int = 64, hi: int = 1024):
    """Log-uniform side lengths on [lo, hi], rounded to integers; (count, 2) as (nx, ny)."""
    u = uniform01(seed, 2 * count, stream=3).reshape(count, 2)
    s = np.exp(math.log(lo) + u * (math.log(hi) - math.log(lo)))
    return np.clip(np.rint(s).astype(np.int64), lo, hi)
