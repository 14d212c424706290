"""torch.autograd binding of the forward and adjoint kernels.

``ensemble_spins(theta, opt)`` maps the 15 field parameters theta (order of
``spintrack.adjoint.GRAD_PARAMS``: Gx, Gy, Gz rows, E, B0) to the
ensemble-mean spin at every output time, a (K, 3) float64 tensor; any loss
built from it with torch operations back-propagates into theta through the
adjoint kernel (one adjoint pass per backward call).

The tensors live on the CPU: the GPU work runs inside libspintrack.so on
its own HIP runtime, which is loaded before torch so that torch binds to the
same one (torch is only used for CPU tensors and autograd here; do not mix
in torch.cuda in the same process).
"""
from __future__ import annotations

import math

import numpy as np

from . import _lib as L

L.load()   # the library's HIP runtime first

import torch  # noqa: E402

from .adjoint import ForwardRun, with_params  # noqa: E402


class EnsembleSpins(torch.autograd.Function):
    """theta (15,) -> <S>(t_k) (K, 3)."""

    @staticmethod
    def forward(ctx, theta, opt, variant, chunk_records):
        run = ForwardRun(with_params(opt, theta.detach().cpu().numpy()), variant=variant,
                         chunk_records=chunk_records)
        try:
            means = run.run()
        except Exception:
            run.close()
            raise
        ctx.run = run
        return torch.from_numpy(means)

    @staticmethod
    def backward(ctx, grad_means):
        run = ctx.run
        try:
            g = run.gradient(grad_means.detach().cpu().numpy())
        finally:
            run.close()
            ctx.run = None
        return torch.from_numpy(g), None, None, None


def ensemble_spins(theta: torch.Tensor, opt, variant: int = L.VARIANT_FAST, chunk_records: int = 256):
    return EnsembleSpins.apply(theta, opt, variant, chunk_records)


def phase_loss(means: torch.Tensor, times, t0: float) -> torch.Tensor:
    """<Sx>(t_K) + accumulated precession phase of the ensemble-mean
    transverse spin (least-squares slope of the unwrapped atan2(<Sz>, <Sy>)
    times the run length); the torch twin of spintrack.adjoint.phase_loss."""
    t = torch.as_tensor(np.asarray(times, dtype=np.float64))
    phi = torch.atan2(means[:, 2], means[:, 1])
    # unwrap: piecewise-constant 2*pi offsets (no gradient through them)
    d = torch.diff(phi.detach())
    jumps = torch.round(d / (2 * math.pi))
    offs = torch.cat([torch.zeros(1, dtype=phi.dtype), -2 * math.pi * torch.cumsum(jumps, 0)])
    phi = phi + offs
    dt = t - t.mean()
    w = torch.sum(dt * (phi - phi.mean())) / torch.sum(dt * dt)
    return means[-1, 0] + w * (t[-1] - t0)
