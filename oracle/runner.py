"""ORACLE -- test infrastructure only.

ctypes drivers for CPU machines: the classes emitted by ``cpp_emitter`` and
the reference's shipped pre-generated classes built by ``oracle/Makefile``.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from paraiso_amd.om.dyn import Realm

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")


class CpuMachine:
    """Drives a compiled CPU machine through its C ABI (prefix ``oracle_``/``ref_``)."""

    def __init__(self, so_path: str, statics, prefix: str = "oracle", memory_shape=None,
                 kernel_names=None):
        self.lib = ctypes.CDLL(so_path)
        self.prefix = prefix
        self.statics = list(statics)
        self.kernel_names = list(kernel_names or [])
        f = self._f
        f("create").restype = ctypes.c_void_p
        f("destroy").argtypes = [ctypes.c_void_p]
        f("run").argtypes = [ctypes.c_void_p, ctypes.c_int]
        f("static_ptr").restype = ctypes.c_void_p
        f("static_ptr").argtypes = [ctypes.c_void_p, ctypes.c_int]
        f("memory_size").restype = ctypes.c_int
        self.h = ctypes.c_void_p(f("create")())
        self.memory_size = f("memory_size")()
        self.memory_shape = tuple(memory_shape) if memory_shape else (self.memory_size,)

    def _f(self, name):
        return getattr(self.lib, "%s_%s" % (self.prefix, name))

    def close(self):
        if self.h:
            self._f("destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _index(self, name_or_idx):
        if isinstance(name_or_idx, int):
            return name_or_idx
        for i, s in enumerate(self.statics):
            if s.name == name_or_idx:
                return i
        raise KeyError(name_or_idx)

    def view(self, name_or_idx) -> np.ndarray:
        """A numpy view of a static (array: memory shape, reversed to C order)."""
        i = self._index(name_or_idx)
        st = self.statics[i]
        ptr = self._f("static_ptr")(self.h, i)
        dt = np.dtype(st.dyn.ctype.np_dtype)
        if st.dyn.realm is Realm.Array:
            n = self.memory_size
            buf = (ctypes.c_char * (n * dt.itemsize)).from_address(ptr)
            arr = np.frombuffer(buf, dtype=dt, count=n)
            return arr.reshape(tuple(reversed(self.memory_shape)))
        buf = (ctypes.c_char * dt.itemsize).from_address(ptr)
        return np.frombuffer(buf, dtype=dt, count=1)

    def get(self, name):
        v = self.view(name)
        st = self.statics[self._index(name)]
        return v.copy() if st.dyn.realm is Realm.Array else v[0].item()

    def set(self, name, value):
        v = self.view(name)
        v[...] = value

    def run(self, kernel, n: int = 1):
        k = kernel if isinstance(kernel, int) else self.kernel_names.index(kernel)
        for _ in range(n):
            self._f("run")(self.h, k)


def emitted_machine(plan, abs_mode: str = "fabs", openmp: bool = True,
                    opt: str = "-O2") -> CpuMachine:
    from . import cpp_emitter
    so = cpp_emitter.build(plan, abs_mode=abs_mode, opt=opt, openmp=openmp)
    return CpuMachine(so, plan.om.setup.static_values, "oracle", plan.memory_size(),
                      [k.name for k in plan.om.kernels])


def reference_machine(which: str, statics, memory_shape, kernel_names) -> CpuMachine:
    so = os.path.join(REF_DIR, "%s_ref.so" % which)
    if not os.path.exists(so):
        raise FileNotFoundError(so + " (run make -C oracle)")
    return CpuMachine(so, statics, "ref", memory_shape, kernel_names)
