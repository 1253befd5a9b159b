"""Python host binding of a generated HIP machine (through ``include/paraiso_hip.h``).

``Machine`` mirrors the generated C++ class: ``run("proceed", n)`` calls an
OM kernel ``n`` times, ``get``/``set`` read and write statics, ``devptr``
returns the device buffer of a static (for zero-copy torch views), and
``tensor`` wraps a static as a torch tensor on ``cuda:0`` without copying.

There is no CPU fallback: constructing a ``Machine`` without the built
libraries or without a GPU raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

from .generator import runtime_library
from .om.dyn import CType

PR_CT = {0: CType.Bool, 1: CType.Int, 2: CType.Float, 3: CType.Double}


class PrStaticDesc(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("realm", ctypes.c_int), ("ctype", ctypes.c_int),
                ("flags", ctypes.c_int)]


class PrLaunchDesc(ctypes.Structure):
    _fields_ = [("fn", ctypes.c_void_p), ("tile", ctypes.c_int * 3), ("n_reduce", ctypes.c_int),
                ("pack", ctypes.c_void_p), ("args_size", ctypes.c_size_t),
                ("batched", ctypes.c_void_p)]


class PrKernelDesc(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("n_launch", ctypes.c_int),
                ("launches", ctypes.POINTER(PrLaunchDesc)), ("n_stored", ctypes.c_int),
                ("stored", ctypes.POINTER(ctypes.c_int)), ("n_loaded", ctypes.c_int),
                ("loaded", ctypes.POINTER(ctypes.c_int)), ("n_written", ctypes.c_int),
                ("written", ctypes.POINTER(ctypes.c_int))]


class PrMachineDesc(ctypes.Structure):
    _fields_ = [("abi", ctypes.c_int), ("name", ctypes.c_char_p), ("ndim", ctypes.c_int),
                ("n_static", ctypes.c_int), ("statics", ctypes.POINTER(PrStaticDesc)),
                ("n_kernel", ctypes.c_int), ("kernels", ctypes.POINTER(PrKernelDesc)),
                ("n_slot", ctypes.c_int), ("halo", ctypes.c_int),
                ("lower_margin", ctypes.c_int * 3), ("upper_margin", ctypes.c_int * 3)]


RUNTIME_SYMBOLS = ["pr_create", "pr_destroy", "pr_set_stream", "pr_sync", "pr_last_error",
                   "pr_run", "pr_run_graph", "pr_run_timed", "pr_static_devptr", "pr_static_write",
                   "pr_static_read", "pr_static_bytes", "pr_size", "pr_memory_size", "pr_cells",
                   "pr_state_save", "pr_state_load",
                   "pr_snapshot_alloc", "pr_snapshot_save", "pr_snapshot_load",
                   "pr_adjoint_forward", "pr_adjoint_backward", "pr_adjoint_forward_seq",
                   "pr_adjoint_backward_seq",
                   "pr_batch_create", "pr_batch_destroy", "pr_batch_grid", "pr_batch_run",
                   "pr_batch_run_graph", "pr_batch_run_timed", "pr_batch_sync",
                   "pr_batch_last_error", "pr_batch_cells"]

_RT = None


def load_runtime(path: Optional[str] = None) -> ctypes.CDLL:
    global _RT
    if _RT is not None and path is None:
        return _RT
    so = path or runtime_library()
    if not os.path.exists(so):
        raise RuntimeError("HIP runtime library %s is not built (run __graft_entry__.build())" % so)
    lib = ctypes.CDLL(so, mode=ctypes.RTLD_GLOBAL)
    vp, i, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t
    lib.pr_create.argtypes = [vp, ctypes.POINTER(i), ctypes.POINTER(vp)]
    lib.pr_destroy.argtypes = [vp]
    lib.pr_set_stream.argtypes = [vp, vp]
    lib.pr_sync.argtypes = [vp]
    lib.pr_last_error.argtypes = [vp]
    lib.pr_last_error.restype = ctypes.c_char_p
    lib.pr_run.argtypes = [vp, i, i]
    lib.pr_run_graph.argtypes = [vp, i, i]
    lib.pr_run_timed.argtypes = [vp, i, i, ctypes.POINTER(ctypes.c_double)]
    lib.pr_static_devptr.argtypes = [vp, i, ctypes.POINTER(vp)]
    lib.pr_static_write.argtypes = [vp, i, vp, sz, i]
    lib.pr_static_read.argtypes = [vp, i, vp, sz, i]
    lib.pr_static_bytes.argtypes = [vp, i]
    lib.pr_static_bytes.restype = sz
    lib.pr_size.argtypes = [vp, i]
    lib.pr_memory_size.argtypes = [vp, i]
    lib.pr_state_save.argtypes = [vp, ctypes.c_char_p]
    lib.pr_state_load.argtypes = [vp, ctypes.c_char_p]
    lib.pr_cells.argtypes = [vp]
    lib.pr_cells.restype = ctypes.c_longlong
    lib.pr_snapshot_alloc.argtypes = [vp, i, ctypes.POINTER(i), i]
    lib.pr_adjoint_forward.argtypes = [vp, i, i, i]
    lib.pr_adjoint_backward.argtypes = [vp, i, i, i, i]
    ip = ctypes.POINTER(i)
    lib.pr_adjoint_forward_seq.argtypes = [vp, ip, i, i, i]
    lib.pr_adjoint_backward_seq.argtypes = [vp, ip, i, i, ip, i, i, i]
    lib.pr_batch_create.argtypes = [vp, i, ctypes.POINTER(i), ctypes.POINTER(vp)]
    lib.pr_batch_destroy.argtypes = [vp]
    lib.pr_batch_grid.argtypes = [vp, i]
    lib.pr_batch_grid.restype = vp
    lib.pr_batch_run.argtypes = [vp, i, i]
    lib.pr_batch_run_graph.argtypes = [vp, i, i]
    lib.pr_batch_run_timed.argtypes = [vp, i, i, ctypes.POINTER(ctypes.c_double)]
    lib.pr_batch_sync.argtypes = [vp]
    lib.pr_batch_last_error.argtypes = [vp]
    lib.pr_batch_last_error.restype = ctypes.c_char_p
    lib.pr_batch_cells.argtypes = [vp]
    lib.pr_batch_cells.restype = ctypes.c_longlong
    lib.pr_snapshot_save.argtypes = [vp, i]
    lib.pr_snapshot_load.argtypes = [vp, i]
    if path is None:
        _RT = lib
    return lib


class PrError(RuntimeError):
    pass


class Machine:
    def __init__(self, machine_so: str, local_size: Sequence[int]):
        if not os.path.exists(machine_so):
            raise RuntimeError("machine library %s is not built" % machine_so)
        self.rt = load_runtime()
        self.lib = ctypes.CDLL(machine_so)
        self.lib.pm_describe.restype = ctypes.POINTER(PrMachineDesc)
        self.desc_ptr = self.lib.pm_describe()
        d = self.desc_ptr.contents
        self.name = d.name.decode()
        self.ndim = d.ndim
        self.statics = []
        for s in range(d.n_static):
            sd = d.statics[s]
            self.statics.append((sd.name.decode(), sd.realm, PR_CT[sd.ctype]))
        self.kernels = [d.kernels[k].name.decode() for k in range(d.n_kernel)]
        n = list(local_size) + [1] * (3 - len(local_size))
        self.local_shape = tuple(local_size)
        # arrays carry the margins of Open axes (om_memory_size_k)
        self.shape = tuple(int(x) + d.lower_margin[a] + d.upper_margin[a]
                           for a, x in enumerate(local_size))
        self.lower_margin = tuple(d.lower_margin[a] for a in range(self.ndim))
        self.upper_margin = tuple(d.upper_margin[a] for a in range(self.ndim))
        arr = (ctypes.c_int * 3)(*n)
        h = ctypes.c_void_p()
        rc = self.rt.pr_create(ctypes.cast(self.desc_ptr, ctypes.c_void_p), arr, ctypes.byref(h))
        if rc != 0:
            raise PrError("pr_create failed (%d) for %s %s" % (rc, self.name, list(local_size)))
        self.h = h

    # ---------------------------------------------------------------- basics
    def _check(self, rc: int, what: str):
        if rc != 0:
            msg = self.rt.pr_last_error(self.h)
            raise PrError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    def close(self):
        if getattr(self, "h", None):
            self.rt.pr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def static_index(self, name) -> int:
        if isinstance(name, int):
            return name
        for i, (nm, _, _) in enumerate(self.statics):
            if nm == name:
                return i
        raise KeyError(name)

    def kernel_index(self, name) -> int:
        return name if isinstance(name, int) else self.kernels.index(name)

    def run(self, kernel, count: int = 1, graph: bool = False):
        k = self.kernel_index(kernel)
        fn = self.rt.pr_run_graph if graph else self.rt.pr_run
        self._check(fn(self.h, k, count), "run %s" % kernel)

    def run_timed(self, kernel, count: int = 1):
        """Run with an event pair around every launch; returns ms per launch index (summed)."""
        k = self.kernel_index(kernel)
        nl = self.desc_ptr.contents.kernels[k].n_launch
        out = (ctypes.c_double * nl)()
        self._check(self.rt.pr_run_timed(self.h, k, count, out), "run_timed %s" % kernel)
        return list(out)

    def sync(self):
        self._check(self.rt.pr_sync(self.h), "sync")

    def set_stream(self, stream_handle: int):
        self._check(self.rt.pr_set_stream(self.h, ctypes.c_void_p(stream_handle)), "set_stream")
        self._stream = stream_handle

    def use_stream(self, stream_handle: int):
        """Issue the machine's work on ``stream_handle`` from now on (set_stream, which waits
        for the current stream and drops cached graphs, only when it differs)."""
        if getattr(self, "_stream", None) != stream_handle:
            self.set_stream(stream_handle)

    # ---------------------------------------------------------------- statics
    def _np_shape(self, s):
        name, realm, ct = self.statics[s]
        if realm == 1:
            return tuple(reversed(self.shape))
        return ()

    def get(self, name):
        s = self.static_index(name)
        _, realm, ct = self.statics[s]
        out = np.empty(self._np_shape(s), dtype=ct.np_dtype)
        self._check(self.rt.pr_static_read(self.h, s, out.ctypes.data_as(ctypes.c_void_p),
                                           out.nbytes, 0), "read %s" % name)
        return out if realm == 1 else out.item()

    def set(self, name, value):
        s = self.static_index(name)
        _, realm, ct = self.statics[s]
        arr = np.ascontiguousarray(np.broadcast_to(np.asarray(value, dtype=ct.np_dtype),
                                                   self._np_shape(s)))
        self._check(self.rt.pr_static_write(self.h, s, arr.ctypes.data_as(ctypes.c_void_p),
                                            arr.nbytes, 0), "write %s" % name)

    def devptr(self, name) -> int:
        s = self.static_index(name)
        p = ctypes.c_void_p()
        self._check(self.rt.pr_static_devptr(self.h, s, ctypes.byref(p)), "devptr")
        return p.value

    def write_device(self, name, src_ptr: int, nbytes: int):
        s = self.static_index(name)
        self._check(self.rt.pr_static_write(self.h, s, ctypes.c_void_p(src_ptr), nbytes, 1),
                    "device write %s" % name)

    def read_device(self, name, dst_ptr: int, nbytes: int):
        s = self.static_index(name)
        self._check(self.rt.pr_static_read(self.h, s, ctypes.c_void_p(dst_ptr), nbytes, 1),
                    "device read %s" % name)

    def nbytes(self, name) -> int:
        return int(self.rt.pr_static_bytes(self.h, self.static_index(name)))

    # ---------------------------------------------------------------- snapshots
    def snapshot_alloc(self, n: int, statics):
        idx = [self.static_index(x) for x in statics]
        arr = (ctypes.c_int * max(1, len(idx)))(*idx)
        self._check(self.rt.pr_snapshot_alloc(self.h, n, arr, len(idx)), "snapshot_alloc")

    def snapshot_save(self, slot: int):
        self._check(self.rt.pr_snapshot_save(self.h, slot), "snapshot_save")

    def snapshot_load(self, slot: int):
        self._check(self.rt.pr_snapshot_load(self.h, slot), "snapshot_load")

    def save(self, path: str):
        """Write every static to `path` (pr_state_save)."""
        self._check(self.rt.pr_state_save(self.h, os.fsencode(path)), "state save")

    def load(self, path: str):
        """Restore the statics from a file written by `save` of the same machine and size."""
        self._check(self.rt.pr_state_load(self.h, os.fsencode(path)), "state load")

    def _kernels(self, ks):
        ks = [ks] if isinstance(ks, (str, int)) else list(ks)
        return (ctypes.c_int * len(ks))(*[self.kernel_index(k) for k in ks]), len(ks)

    def adjoint_forward(self, fwd, nsteps: int, every: int):
        """``nsteps`` steps of ``fwd`` (a kernel or a sequence of kernels), a checkpoint
        every ``every`` steps (pr_adjoint_forward_seq)."""
        f, nf = self._kernels(fwd)
        self._check(self.rt.pr_adjoint_forward_seq(self.h, f, nf, nsteps, every),
                    "adjoint_forward")

    def adjoint_backward(self, fwd, adj, nsteps: int, every: int, tape_after: int = 0):
        """The checkpointed adjoint (pr_adjoint_backward_seq): ``fwd`` and ``adj`` are a
        kernel or a sequence; the adjoint sees the state after ``fwd[:tape_after]``."""
        f, nf = self._kernels(fwd)
        a, na = self._kernels(adj)
        self._check(self.rt.pr_adjoint_backward_seq(self.h, f, nf, tape_after, a, na, nsteps,
                                                    every), "adjoint_backward")


class _GridView(Machine):
    """A grid of a Batch: a Machine that does not own its handle."""

    def __init__(self, batch, g):   # noqa: D401 - no Machine.__init__
        self.rt, self.lib, self.desc_ptr = batch.rt, batch.lib, batch.desc_ptr
        self.name, self.ndim, self.statics, self.kernels = (batch.name, batch.ndim, batch.statics,
                                                            batch.kernels)
        self.h = ctypes.c_void_p(batch.rt.pr_batch_grid(batch.b, g))
        self.local_shape = tuple(batch.sizes[g])
        self.shape = tuple(self.rt.pr_memory_size(self.h, a) for a in range(self.ndim))
        self._owner = batch

    def close(self):
        self.h = None


class Batch:
    """Many grids of one machine, of uneven sizes, advanced by the same launches."""

    def __init__(self, machine_so: str, sizes):
        if not os.path.exists(machine_so):
            raise RuntimeError("machine library %s is not built" % machine_so)
        self.rt = load_runtime()
        self.lib = ctypes.CDLL(machine_so)
        self.lib.pm_describe.restype = ctypes.POINTER(PrMachineDesc)
        self.desc_ptr = self.lib.pm_describe()
        d = self.desc_ptr.contents
        self.name, self.ndim = d.name.decode(), d.ndim
        self.statics = [(d.statics[s].name.decode(), d.statics[s].realm, PR_CT[d.statics[s].ctype])
                        for s in range(d.n_static)]
        self.kernels = [d.kernels[k].name.decode() for k in range(d.n_kernel)]
        self.sizes = [tuple(int(x) for x in sz) for sz in sizes]
        flat = []
        for sz in self.sizes:
            flat += list(sz) + [1] * (3 - len(sz))
        arr = (ctypes.c_int * len(flat))(*flat)
        b = ctypes.c_void_p()
        rc = self.rt.pr_batch_create(ctypes.cast(self.desc_ptr, ctypes.c_void_p), len(self.sizes),
                                     arr, ctypes.byref(b))
        if rc != 0:
            raise PrError("pr_batch_create failed (%d)" % rc)
        self.b = b
        self.grids = [_GridView(self, g) for g in range(len(self.sizes))]

    def _check(self, rc, what):
        if rc != 0:
            msg = self.rt.pr_batch_last_error(self.b)
            raise PrError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    def run(self, kernel, count=1, graph=False):
        k = self.kernels.index(kernel) if isinstance(kernel, str) else kernel
        fn = self.rt.pr_batch_run_graph if graph else self.rt.pr_batch_run
        self._check(fn(self.b, k, count), "batch run %s" % kernel)

    def run_timed(self, kernel, count=1):
        k = self.kernels.index(kernel) if isinstance(kernel, str) else kernel
        nl = self.desc_ptr.contents.kernels[k].n_launch
        out = (ctypes.c_double * nl)()
        self._check(self.rt.pr_batch_run_timed(self.b, k, count, out), "batch run_timed")
        return list(out)

    def sync(self):
        self._check(self.rt.pr_batch_sync(self.b), "batch sync")

    def cells(self) -> int:
        return int(self.rt.pr_batch_cells(self.b))

    def close(self):
        if getattr(self, "b", None):
            self.rt.pr_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
