"""Code-generation setup (``Language/Paraiso/Generator/Native.hs:16-43``).

The reference selects ``CPlusPlus`` or ``CUDA``; this backend adds ``HIP``,
the MI355X path.  ``cuda_grid_size`` is kept for source compatibility and is
read as a launch hint (workgroup tile) by the HIP generator.
"""
from __future__ import annotations

from dataclasses import dataclass, replace
from enum import Enum
from typing import Optional, Tuple

from .. import annotation as A


class Language(Enum):
    CPlusPlus = "CPlusPlus"
    CUDA = "CUDA"
    HIP = "HIP"


@dataclass
class Setup:
    local_size: Tuple[int, ...]
    language: Language = Language.HIP
    directory: str = "./"
    opt_level: A.Level = A.Level.O3
    boundary: Tuple[A.Condition, ...] = None
    cuda_grid_size: Tuple[int, int] = (32, 32)
    # HIP-only knobs (tile shape, temporal unroll); None lets the generator choose
    tile: Optional[Tuple[int, ...]] = None
    steps_per_launch: int = 1
    batched: bool = False        # also emit the many-grids form of every launch
    threads: int = 256           # work-items per workgroup (a multiple of 64, at most 1024)
    # per-point operations a value needed at several offsets must save to be computed once
    # into LDS instead of per offset in registers: the Manifest/Delayed choice of
    # DecideAllocation.hs:41-43 (AllocationChoice), as a tuning knob
    mat_gain: int = 48
    # > 0: each workgroup streams along the last axis through its tile (whose last extent
    # is then the streaming length), this many planes per step
    march: int = 0
    # with march: load the next step's planes into registers while the current step computes
    prefetch: bool = True
    # block tiles: launch as many workgroups as fit on the device at once and let each loop
    # over tiles, loading the next tile's fields while it computes the current one
    persist: bool = False
    # after the greedy LDS choice, also compute in an earlier phase, as one group, the values
    # a later phase evaluates in registers from mostly the same expansion, and let a launch
    # take the whole CU's LDS (for launches whose registers allow one workgroup per CU)
    join: bool = False
    # per OM kernel, a launch form overriding the above: ((kernel name, {"tile": ...,
    # "threads": ..., "persist": ..., "march": ..., "mat_gain": ..., "join": ...}), ...)
    kernel_forms: Tuple[Tuple[str, dict], ...] = ()

    def __post_init__(self):
        self.local_size = tuple(int(x) for x in self.local_size)
        if self.boundary is None:
            self.boundary = tuple(A.Condition.Open for _ in self.local_size)
        self.boundary = tuple(self.boundary)
        if len(self.boundary) != len(self.local_size):
            raise ValueError("boundary and local_size differ in dimension")
        if self.threads % 64 or not 64 <= self.threads <= 1024:
            raise ValueError("threads must be a multiple of 64 in [64, 1024]")

    @property
    def ndim(self) -> int:
        return len(self.local_size)

    def with_(self, **kw) -> "Setup":
        return replace(self, **kw)

    def form_of(self, kernel: str) -> dict:
        """The launch form of one OM kernel: the machine's, with the kernel's overrides."""
        f = {"tile": self.tile, "threads": self.threads, "persist": self.persist,
             "march": self.march, "mat_gain": self.mat_gain, "join": self.join}
        for name, over in self.kernel_forms:
            if name == kernel:
                f.update(over)
        return f


def default_setup(local_size) -> Setup:
    """``defaultSetup`` (Native.hs:98-107): Open boundaries, O3, grid (32,32)."""
    return Setup(local_size=tuple(local_size), language=Language.HIP)
