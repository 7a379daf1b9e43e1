"""Dual-energy elemental decomposition by gradient descent on the differentiable projector.

Several problems (phantoms of different slice counts) are optimised together: each Adam iteration runs, for
every problem and every spectrum, the polychromatic log projection and its adjoint (the per-energy line
integrals are recomputed by the adjoint kernel, never stored).
"""
from __future__ import annotations

import torch

from .autograd import poly_logproj


def decompose(problems, iters=200, lr=0.01, init=None, nonneg=True, callback=None):
    """problems: list of dicts {"projectors": [Projector per spectrum, same geometry], "targets": [p per
    spectrum, FP32 device tensors], "shape": (6, Z, N, N), "xfactor": float}.  Returns the list of fitted
    maps (FP32 [mat][z][y][x] device tensors) and the loss history."""
    maps = []
    for i, pr in enumerate(problems):
        m = (init[i].clone() if init is not None else
             torch.zeros(pr["shape"], dtype=torch.float32, device=pr["targets"][0].device))
        maps.append(m.requires_grad_(True))
    opt = torch.optim.Adam(maps, lr=lr)
    hist = []
    for it in range(iters):
        opt.zero_grad(set_to_none=True)
        total = None
        for m, pr in zip(maps, problems):
            for P, t in zip(pr["projectors"], pr["targets"]):
                p = poly_logproj(m, P, pr.get("xfactor", 1.0))
                loss = ((p.double() - t.double()) ** 2).sum()
                loss.backward()
                total = loss.detach() if total is None else total + loss.detach()
        opt.step()
        if nonneg:
            with torch.no_grad():
                for m in maps:
                    m.clamp_(min=0.0)
        hist.append(total)
        if callback:
            callback(it, total)
    hist = [float(h) for h in hist]  # one synchronisation at the end
    return [m.detach() for m in maps], hist
