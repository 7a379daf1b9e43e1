"""Detector model between the forward projector and FDK, on the device: scatter composition and quantum
noise.  No reference counterpart: the reference leaves this stage to other programs (its projection file
names, F/txtfile; the scatter U-Net, Unet_scatter.py:349-423, predicts the scatter that is added here).

    noisy = detector_counts(primary, scatter=None, I0=1e5, seed=0)   # max(k, 1) / I0 per pixel
    counts = detector_counts(primary, I0=1e5, raw=True)              # k itself

k ~ Poisson(I0 * (primary + scatter)), drawn from Philox4x32-10 keyed by (seed, flat pixel index): the draw
of a pixel does not depend on the launch or on the other pixels.  The normalised output feeds
fdk.filter_sinogram_device (which forms -log p) directly.
"""
from __future__ import annotations

import ctypes as C

from ._lib import check, lib
from ._lib import stream_arg as _stream


def _torch():
    import torch
    return torch


def detector_counts(primary, scatter=None, I0=1.0e5, seed=0, raw=False, out=None, stream=None):
    """Poisson counts of I0 (primary + scatter) per pixel of contiguous float32 device tensors of one shape:
    k (raw) or max(k, 1) / I0.  `out` may be `primary` (in place)."""
    torch = _torch()

    def _check(t, name):
        if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous():
            raise ValueError(f"{name} must be a contiguous float32 CUDA tensor")

    _check(primary, "primary")
    if scatter is not None:
        _check(scatter, "scatter")
        if scatter.shape != primary.shape:
            raise ValueError("scatter must have the primary's shape")
    if out is None:
        out = torch.empty_like(primary)
    else:
        _check(out, "out")
        if out.shape != primary.shape:
            raise ValueError("out must have the primary's shape")
    if not 0 <= int(seed) < 1 << 64:
        raise ValueError("seed must fit in 64 bits")
    check(lib().vcbct_detector_counts(primary.numel(), C.c_void_p(primary.data_ptr()),
                                      None if scatter is None else C.c_void_p(scatter.data_ptr()), float(I0),
                                      int(seed), int(bool(raw)), C.c_void_p(out.data_ptr()), _stream(stream)))
    return out
