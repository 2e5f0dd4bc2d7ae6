"""torch.autograd binding of the differentiable rollout.

``rollout(engine, fields, pos, vel, ene, steps, dt)`` returns the final
(pos, vel, ene) after Solver::initialize and ``steps`` fixed-``dt`` steps on
the GPU; its backward calls ``sphgpu_rollout_adjoint`` of the same context.
Inputs may live on any device: the C ABI takes host buffers, so they are
copied through host memory (the rollout itself runs entirely on the GPU).

A context holds the checkpoints of its last rollout only: run the backward
of one rollout before the next forward on the same engine, or use one
engine per pending graph.
"""
from __future__ import annotations

import itertools

import numpy as np
import torch

_generation = itertools.count(1)


def _np(t: torch.Tensor) -> np.ndarray:
    return t.detach().to("cpu", torch.float64).contiguous().numpy()


class Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, vel, ene, engine, fields, steps, dt):
        f = dict(fields)
        f["pos"], f["vel"], f["ene"] = _np(pos), _np(vel), _np(ene)
        engine.upload(f)
        engine.rollout(int(steps), float(dt))
        out = engine.download(("pos", "vel", "ene"))
        ctx.engine = engine
        ctx.generation = next(_generation)
        engine._rollout_generation = ctx.generation
        dev, dt_ = pos.device, pos.dtype
        return (torch.from_numpy(out["pos"]).to(dev, dt_), torch.from_numpy(out["vel"]).to(dev, dt_),
                torch.from_numpy(out["ene"]).to(dev, dt_))

    @staticmethod
    def backward(ctx, gx, gv, gu):
        engine = ctx.engine
        if getattr(engine, "_rollout_generation", None) != ctx.generation:
            raise RuntimeError("the engine ran another rollout since this forward; its checkpoints are gone")
        dx, dv, du = engine.rollout_adjoint(None if gx is None else _np(gx), None if gv is None else _np(gv),
                                            None if gu is None else _np(gu))
        ref = gx if gx is not None else gv if gv is not None else gu
        dev, dt_ = ref.device, ref.dtype
        return (torch.from_numpy(dx).to(dev, dt_), torch.from_numpy(dv).to(dev, dt_),
                torch.from_numpy(du).to(dev, dt_), None, None, None, None)


def rollout(engine, fields: dict, pos, vel, ene, steps: int, dt: float):
    """Differentiable K-step rollout; ``fields`` supplies mass and the other
    initial fields, ``pos``/``vel``/``ene`` are the differentiated inputs."""
    return Rollout.apply(pos, vel, ene, engine, fields, steps, dt)
