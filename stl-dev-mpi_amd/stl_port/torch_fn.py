"""torch.autograd wrappers of NativeEngine (reference
examples/06_differentiable_python/stl_torch.py:25-42 and
amortized/stl_torch.py:29-53), on device memory: the maps and the gradient
stay on the GPU (stlp_forward_device / stlp_vjp_device); only the
coefficients and their cotangents cross to the host.

Complex outputs are complex128 tensors; torch hands their backward the
cotangent dL/dRe + i dL/dIm, which is the engine's vjp convention.  An engine
keeps the state of its last forward only: run backward before the next
forward of the same engine.
"""
import itertools

import numpy as np
import torch

_generation = itertools.count(1)


class _Scattering(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, engine):
        g = next(_generation)
        xs = x.detach().to(torch.float64).contiguous()
        values, _ = engine.forward_device(xs.reshape(engine._map_shape()), generation=g)
        ctx.engine, ctx.generation, ctx.shape = engine, g, x.shape
        return torch.from_numpy(values).to(x.device)

    @staticmethod
    def backward(ctx, gy):
        cot = gy.detach().to(torch.complex128).cpu().numpy()
        out = torch.empty(ctx.engine._map_shape(), dtype=torch.float64, device=gy.device)
        ctx.engine.vjp_device(cot.reshape(ctx.engine._shape()), out, generation=ctx.generation)
        return out.reshape(ctx.shape), None


def scattering(engine, x):
    """Statistics of NativeEngine as a torch op: x [B, npix] (device) -> complex128 [B, ncoeff]
    ([B, K, ncoeff] with K masks)."""
    return _Scattering.apply(x, engine)


class StatsFunction(torch.autograd.Function):
    """Real view [B, ncoeff, 2] of the statistics (amortized/stl_torch.py:29-41)."""

    @staticmethod
    def forward(ctx, x, engine, generation):
        xs = x.detach().to(torch.float64).contiguous()
        values, _ = engine.forward_device(xs, generation=generation)
        ctx.engine, ctx.generation = engine, generation
        out = values.view(np.float64).reshape(values.shape + (2,))
        return torch.from_numpy(out.copy()).to(x.device)

    @staticmethod
    def backward(ctx, g):
        device = g.device
        gh = g.detach().to(torch.float64).cpu().numpy()
        out = torch.empty(ctx.engine._map_shape(), dtype=torch.float64, device=device)
        ctx.engine.vjp_device(gh[..., 0] + 1j * gh[..., 1], out, generation=ctx.generation)
        return out, None, None


class Stats:
    """Statistics of a batch with gradient; one engine, one forward state at a time."""

    def __init__(self, engine):
        self.engine = engine
        self.generation = 0

    def __call__(self, x):
        self.generation += 1
        return StatsFunction.apply(x, self.engine, self.generation)
