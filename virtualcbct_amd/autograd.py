"""Differentiable polychromatic projector for PyTorch.

`poly_logproj(maps, projector, xfactor)` returns the polychromatic log projection
    p = -log( sum_e pdf_e exp(-sum_m mu_me L_m / xfactor) / sum_e pdf_e )
of six FP32 elemental density maps [mat][z][y][x] (a torch tensor on the device).  The forward keeps the six
per-ray material line integrals (24 B per ray; the per-energy terms are never stored, the backward recomputes
them from the line integrals and the projector's spectrum, which must not change in between), and the backward runs
the hand-written adjoint kernels: from them, the per-ray coefficients dL/dp * mubar_m / xfactor (mubar_m the
transmission-weighted spectral mean of mu_m), scattered along a second traversal of the rays.  There is no reference counterpart; the reference's scatter U-Net (Unet_scatter.py)
consumes projections of this kind.
"""
from __future__ import annotations

import torch

from .projector import Projector


class PolyLogProjection(torch.autograd.Function):
    @staticmethod
    def forward(ctx, maps, projector: Projector, xfactor: float):
        if maps.dtype != torch.float32 or not maps.is_cuda:
            raise TypeError("maps must be an FP32 tensor on the device")
        m = maps.contiguous()
        projector.set_volume(m)
        p, L = projector.logproj_save(xfactor)
        ctx.projector = projector
        ctx.xfactor = xfactor
        ctx.spectrum_generation = projector.spectrum_generation
        ctx.save_for_backward(L)
        return p

    @staticmethod
    def backward(ctx, grad_p):
        (L,) = ctx.saved_tensors
        # the backward reads the geometry and the projector's current spectrum, not the maps: the projector may
        # have projected other maps since forward, but its spectrum must still be the one forward used
        if ctx.projector.spectrum_generation != ctx.spectrum_generation:
            raise RuntimeError("poly_logproj: the projector's spectrum was replaced between forward and backward; "
                               "use one projector per spectrum")
        grad = ctx.projector.logproj_backward_saved(L, grad_p.contiguous(), ctx.xfactor)
        return grad, None, None


def poly_logproj(maps, projector: Projector, xfactor: float = 1.0):
    return PolyLogProjection.apply(maps, projector, xfactor)


class DifferentiableProjector(torch.nn.Module):
    """nn.Module wrapper: geometry and spectrum fixed at construction, maps as the input."""

    def __init__(self, projector: Projector, xfactor: float = 1.0):
        super().__init__()
        self.projector = projector
        self.xfactor = float(xfactor)

    def forward(self, maps):
        return poly_logproj(maps, self.projector, self.xfactor)
