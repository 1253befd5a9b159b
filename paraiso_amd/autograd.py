"""torch.autograd bridge for machines that carry an adjoint kernel.

``Rollout.apply(machine, nsteps, every, *fields)`` advances the machine
``nsteps`` calls of ``proceed`` from the given initial fields (CUDA tensors,
one per differentiated static) and returns the final fields.  Its backward
runs the machine's checkpointed adjoint: the forward saved a checkpoint every
``every`` steps (``pr_adjoint_forward``); the backward replays each segment
and applies ``proceed_adjoint`` in reverse (``pr_adjoint_backward``), so the
gradient of any torch function of the final fields with respect to the
initial fields costs one extra forward plus one adjoint step per step, with
ceil(nsteps/every) + every states of storage.

Tensors are exchanged with device-to-device copies on the machine's stream,
which is first synchronised with torch's current stream.  Other statics
(dt, dR, ...) must be set on the machine beforehand.
"""
from __future__ import annotations

from typing import List, Sequence

import torch

from .machine import Machine


def _sync_with_torch(machine: Machine):
    torch.cuda.current_stream().synchronize()


def _write(machine: Machine, name: str, t: torch.Tensor):
    t = t.detach().contiguous()
    machine.write_device(name, t.data_ptr(), t.numel() * t.element_size())


def _read(machine: Machine, name: str, like: torch.Tensor) -> torch.Tensor:
    out = torch.empty_like(like, memory_format=torch.contiguous_format)
    machine.read_device(name, out.data_ptr(), out.numel() * out.element_size())
    return out


class Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, machine: Machine, names: Sequence[str], nsteps: int, every: int,
                fwd: str, adj: str, *fields):
        _sync_with_torch(machine)
        snap = [s for s, (_, _, _) in zip(range(len(machine.statics)), machine.statics)
                if not machine.statics[s][0].startswith("adj_")
                and not machine.statics[s][0].startswith("__scratch")]
        S = (nsteps + every - 1) // every
        machine.snapshot_alloc(S + every, snap)
        for n, f in zip(names, fields):
            _write(machine, n, f)
        machine.adjoint_forward(fwd, nsteps, every)
        outs = tuple(_read(machine, n, f) for n, f in zip(names, fields))
        machine.sync()
        ctx.machine, ctx.names, ctx.nsteps, ctx.every = machine, list(names), nsteps, every
        ctx.fwd, ctx.adj = fwd, adj
        ctx.like = [f.detach() for f in fields]
        return outs

    @staticmethod
    def backward(ctx, *grads):
        m = ctx.machine
        _sync_with_torch(m)
        for n, g, like in zip(ctx.names, grads, ctx.like):
            g = torch.zeros_like(like) if g is None else g
            _write(m, "adj_" + n, g)
        m.adjoint_backward(ctx.fwd, ctx.adj, ctx.nsteps, ctx.every)
        gin = tuple(_read(m, "adj_" + n, like) for n, like in zip(ctx.names, ctx.like))
        m.sync()
        return (None, None, None, None, None, None) + gin


def rollout(machine: Machine, names: Sequence[str], fields: Sequence[torch.Tensor], nsteps: int,
            every: int = 20, fwd: str = "proceed", adj: str = "proceed_adjoint") -> List[torch.Tensor]:
    """Differentiable ``nsteps`` calls of ``fwd`` on ``fields`` (see module docstring)."""
    return list(Rollout.apply(machine, list(names), nsteps, every, fwd, adj, *fields))
