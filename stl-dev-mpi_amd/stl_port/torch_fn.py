"""NativeEngine as a differentiable torch operation, on device memory.

The maps go to the engine as float64 device tensors (stlp_forward_device) and the map gradient
comes back into a device tensor (stlp_vjp_device); only the statistics and their cotangents
cross to the host.  The reference's torch wrappers (examples/06_differentiable_python/
stl_torch.py:25-42, amortized/stl_torch.py:29-53) copy the maps to numpy instead.

Two output forms:
- `scattering(engine, x)`: complex128 statistics [B, (K,) ncoeff].  torch hands the backward of
  a complex output the cotangent dL/dRe + i dL/dIm, which is the engine's vjp convention.
- `StatsFunction` / `Stats`: the statistics as a real tensor [B, ncoeff, 2] (real, imaginary),
  the form the reference's amortized training uses.

An engine holds the state of its last forward only, so a backward must run before the next
forward through the same engine; each forward gets a fresh generation number, and a stale
backward is refused by the engine's identity check.
"""
import itertools

import torch

_generations = itertools.count(1)


def _run_forward(engine, x, generation):
    """Statistics (complex numpy) of x through the engine's device entry point."""
    maps = x.detach().to(dtype=torch.float64).contiguous().reshape(engine._map_shape())
    values, _ = engine.forward_device(maps, generation=generation)
    return values


def _run_vjp(engine, cot, generation, device):
    """Map gradient [engine map shape] of the complex cotangents (numpy) on `device`."""
    grad = torch.empty(engine._map_shape(), dtype=torch.float64, device=device)
    engine.vjp_device(cot.reshape(engine._shape()), grad, generation=generation)
    return grad


class _ComplexStats(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, engine):
        generation = next(_generations)
        values = _run_forward(engine, x, generation)
        ctx.engine, ctx.generation, ctx.in_shape = engine, generation, x.shape
        return torch.from_numpy(values).to(x.device)

    @staticmethod
    def backward(ctx, grad_out):
        cot = grad_out.detach().to(torch.complex128).cpu().numpy()
        grad = _run_vjp(ctx.engine, cot, ctx.generation, grad_out.device)
        return grad.reshape(ctx.in_shape), None


def scattering(engine, x):
    """Complex statistics of x [B, npix] ([B, 2, npix] for cross pairs) on the device, with
    gradient: complex128 [B, ncoeff], or [B, K, ncoeff] with K masks."""
    return _ComplexStats.apply(x, engine)


class StatsFunction(torch.autograd.Function):
    """Real-pair statistics [B, ncoeff, 2] of x with an explicit generation number."""

    @staticmethod
    def forward(ctx, x, engine, generation):
        values = _run_forward(engine, x, generation)
        ctx.engine, ctx.generation, ctx.in_shape = engine, generation, x.shape
        pairs = torch.view_as_real(torch.from_numpy(values))
        return pairs.to(x.device)

    @staticmethod
    def backward(ctx, grad_pairs):
        g = grad_pairs.detach().to(torch.float64).cpu()
        cot = torch.view_as_complex(g.contiguous()).numpy()
        grad = _run_vjp(ctx.engine, cot, ctx.generation, grad_pairs.device)
        return grad.reshape(ctx.in_shape), None, None


class Stats:
    """Callable layer: stats(x) -> StatsFunction of x through one engine, numbering its
    evaluations itself.  One forward state per engine: differentiate before the next call."""

    def __init__(self, engine):
        self.engine = engine
        self.generation = 0

    def __call__(self, x):
        self.generation = next(_generations)
        return StatsFunction.apply(x, self.engine, self.generation)
