"""pr_run_noalloc (the entry point that allocates nothing) against pr_run on a machine that
pr_create allocated: the same kernels on the same input give the same bits, on two calls in
a row, the second on a smaller grid (the library's fixed scratch array keeps the first
call's contents)."""
import ctypes

import numpy as np
import pytest
import torch

from paraiso_amd import inputs, variants
from paraiso_amd.machine import load_runtime

pytestmark = pytest.mark.gpu
FIELDS = ("density", "velocity0", "velocity1", "pressure")


def _bind(rt):
    vp, i = ctypes.c_void_p, ctypes.c_int
    rt.pr_run_noalloc.argtypes = [vp, ctypes.POINTER(i), i, i, ctypes.POINTER(vp),
                                  ctypes.POINTER(i), vp, vp, ctypes.c_char_p, ctypes.c_size_t]
    return rt


def _scalar_block(statics, values):
    """8 bytes per static, scalar static s at byte 8 s (pr_static_devptr's layout)."""
    blk = np.zeros(8 * len(statics), dtype=np.uint8)
    for s, (name, realm, ct) in enumerate(statics):
        if realm == 1 or name not in values:
            continue
        v = np.asarray(values[name], dtype=ct.np_dtype)
        blk[8 * s:8 * s + v.itemsize] = np.frombuffer(v.tobytes(), dtype=np.uint8)
    return blk


def _read_scalar(blk, statics, name):
    s = [x[0] for x in statics].index(name)
    ct = statics[s][2]
    return np.frombuffer(blk[8 * s:8 * s + np.dtype(ct.np_dtype).itemsize].tobytes(),
                         dtype=ct.np_dtype)[0]


def _noalloc_run(rt, ref, nx, ny, seed, steps):
    """KH state on caller-owned torch buffers, proceed_prime + `steps` x proceed_pipelined."""
    statics = ref.statics
    ns = len(statics)
    state = dict(zip(FIELDS, inputs.kh_init(nx, ny, seed, dtype=np.float32)))
    bufs, keep = (ctypes.c_void_p * (2 * ns))(), {}
    for s, (name, realm, ct) in enumerate(statics):
        if realm != 1 or name.startswith("__scratch"):
            continue
        keep[s] = [torch.zeros((ny, nx), dtype=torch.float32, device="cuda") for _ in range(2)]
        if name in state:
            keep[s][0].copy_(torch.from_numpy(state[name]))
        for b in range(2):
            bufs[2 * s + b] = keep[s][b].data_ptr()
    parity = (ctypes.c_int * ns)()
    scal = torch.from_numpy(_scalar_block(statics, inputs.hydro_scalars(nx, ny))).cuda()
    torch.cuda.synchronize()
    err = ctypes.create_string_buffer(512)
    size = (ctypes.c_int * 3)(nx, ny, 1)
    desc = ctypes.cast(ref.desc_ptr, ctypes.c_void_p)
    for kernel, count in (("proceed_prime", 1), ("proceed_pipelined", steps)):
        rc = rt.pr_run_noalloc(desc, size, ref.kernel_index(kernel), count, bufs, parity,
                               ctypes.c_void_p(scal.data_ptr()), None, err, len(err))
        assert rc == 0, (kernel, rc, err.value.decode())
    torch.cuda.synchronize()
    out = {name: keep[s][parity[s]].cpu().numpy()
           for s, (name, _, _) in enumerate(statics) if name in state}
    out["time"] = _read_scalar(scal.cpu().numpy(), statics, "time")
    return out


def test_noalloc_matches_pr_run_twice(require_gpu):
    rt = _bind(load_runtime())
    steps = 12
    for nx, ny, seed in ((320, 192, 3), (96, 64, 4)):      # the second call: a smaller grid
        print("noalloc hydro_f32 %dx%d" % (nx, ny), flush=True)
        ref = variants.machine("hydro_f32", (nx, ny))
        for k, v in inputs.hydro_scalars(nx, ny).items():
            ref.set(k, v)
        for f, a in zip(FIELDS, inputs.kh_init(nx, ny, seed, dtype=np.float32)):
            ref.set(f, a)
        ref.run("proceed_prime")
        ref.run("proceed_pipelined", steps)
        got = _noalloc_run(rt, ref, nx, ny, seed, steps)
        assert got["time"] == np.float32(ref.get("time"))
        for f in FIELDS:
            assert np.array_equal(got[f], ref.get(f)), (nx, ny, f)
        ref.close()


def test_noalloc_rejects_oversize(require_gpu):
    rt = _bind(load_runtime())
    ref = variants.machine("hydro_adj_f32_fast", (64, 64))
    ns = len(ref.statics)
    err = ctypes.create_string_buffer(512)
    size = (ctypes.c_int * 3)(8192, 8192, 1)
    bufs = (ctypes.c_void_p * (2 * ns))()
    parity = (ctypes.c_int * ns)()
    rc = rt.pr_run_noalloc(ctypes.cast(ref.desc_ptr, ctypes.c_void_p), size,
                           ref.kernel_index("proceed"), 1, bufs, parity, None, None, err,
                           len(err))
    assert rc == -1 and b"scratch" in err.value, err.value
