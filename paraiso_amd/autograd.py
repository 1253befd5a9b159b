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

The machine is moved onto torch's current stream (``pr_set_stream``) before it
touches a tensor, so its copies and kernels are ordered with torch's work by the
stream itself: no host synchronisation, and the caching allocator's stream
semantics hold for every temporary.  Other statics (dt, dR, ...) must be set on
the machine beforehand.

The checkpoints of a rollout live in the machine's snapshot slots, so the
machine holds those of its latest rollout only: backward of an earlier rollout
of the same machine raises instead of replaying another run's states.
"""
from __future__ import annotations

from typing import List, Sequence

import torch

from .machine import Machine


def _on_torch_stream(machine: Machine):
    machine.use_stream(torch.cuda.current_stream().cuda_stream)


def _write(machine: Machine, name: str, t: torch.Tensor):
    """Copy ``t`` into a static on the machine's stream, which is torch's current stream:
    the contiguous temporary is freed back to the caching allocator on that same stream,
    after the copy."""
    t = t.detach().contiguous()
    if t.numel() * t.element_size() != machine.nbytes(name):
        raise ValueError("tensor for %s has %d bytes, the static %d"
                         % (name, t.numel() * t.element_size(), machine.nbytes(name)))
    machine.write_device(name, t.data_ptr(), t.numel() * t.element_size())


def _read(machine: Machine, name: str, like: torch.Tensor) -> torch.Tensor:
    out = torch.empty_like(like, memory_format=torch.contiguous_format)
    machine.read_device(name, out.data_ptr(), out.numel() * out.element_size())
    return out


class Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, machine: Machine, names: Sequence[str], nsteps: int, every: int,
                fwd, adj, tape_after: int, *fields):
        _on_torch_stream(machine)
        machine._rollout_gen = getattr(machine, "_rollout_gen", 0) + 1
        ctx.gen = machine._rollout_gen
        snap = [s for s, (_, _, _) in zip(range(len(machine.statics)), machine.statics)
                if not machine.statics[s][0].startswith("adj_")
                and not machine.statics[s][0].startswith("__scratch")]
        S = (nsteps + every - 1) // every
        machine.snapshot_alloc(S + every, snap)
        for n, f in zip(names, fields):
            _write(machine, n, f)
        machine.adjoint_forward(fwd, nsteps, every)
        outs = tuple(_read(machine, n, f) for n, f in zip(names, fields))
        ctx.machine, ctx.names, ctx.nsteps, ctx.every = machine, list(names), nsteps, every
        ctx.fwd, ctx.adj, ctx.tape_after = fwd, adj, tape_after
        ctx.like = [f.detach() for f in fields]
        return outs

    @staticmethod
    def backward(ctx, *grads):
        m = ctx.machine
        if getattr(m, "_rollout_gen", None) != ctx.gen:
            raise RuntimeError("the checkpoints of this rollout were overwritten by a later "
                               "rollout of the same machine: call backward before rolling it "
                               "out again, or use one machine per rollout")
        _on_torch_stream(m)
        for n, g, like in zip(ctx.names, grads, ctx.like):
            g = torch.zeros_like(like) if g is None else g
            _write(m, "adj_" + n, g)
        m.adjoint_backward(ctx.fwd, ctx.adj, ctx.nsteps, ctx.every, ctx.tape_after)
        gin = tuple(_read(m, "adj_" + n, like) for n, like in zip(ctx.names, ctx.like))
        return (None, None, None, None, None, None, None) + gin


def rollout(machine: Machine, names: Sequence[str], fields: Sequence[torch.Tensor], nsteps: int,
            every: int = 20, fwd="proceed", adj="proceed_adjoint",
            tape_after: int = 0) -> List[torch.Tensor]:
    """Differentiable ``nsteps`` steps of ``fwd`` on ``fields`` (see module docstring).
    ``fwd`` and ``adj`` are an OM kernel or a sequence of them (a step split in stages,
    with one adjoint kernel per stage run in reverse order); the adjoint kernels see the
    state after ``fwd[:tape_after]``."""
    return list(Rollout.apply(machine, list(names), nsteps, every, fwd, adj, tape_after,
                              *fields))
