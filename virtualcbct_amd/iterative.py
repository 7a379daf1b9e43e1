"""Iterative-reconstruction pieces of the reference's projector tree, on the device:

  lengthproj    P/lengthproj.cpp:9-131     voxel-driven, chord-weighted backprojection of (65535 - projection)
  IR_ImageCBCT  P/IR_ImageCBCT.cpp:19-94   one MLEM update with a 13-neighbour total-variation penalty

The reference's main program includes both but no longer calls them; they keep its argument meaning
(the projector's geometry for lengthproj, a [Nz][N][N] int16 image for the update) and give its results
bit for bit.
"""
from __future__ import annotations

import ctypes as C

from ._lib import check, lib


def _torch():
    import torch
    return torch


def _stream(stream):
    if stream is not None:
        return C.c_void_p(stream)
    return C.c_void_p(_torch().cuda.current_stream().cuda_stream)


def lengthproj(projector, proj, ite=0, a0=0, a1=None, out=None, stream=None):
    """npf [Nz][N][N] FP32 from proj [n_angles][H][W] (16-bit device tensor of uint16 pixel values) over the
    angles [a0, a1).  ite != 0 resets the registration shifts, as the reference does."""
    torch = _torch()
    g = projector.g
    if proj.element_size() != 2 or not proj.is_contiguous() or proj.numel() != g.n_angles * g.proj_height * g.proj_width:
        raise ValueError("proj must be a contiguous 16-bit device tensor [n_angles][H][W]")
    if out is None:
        out = torch.empty((g.recon_slices, g.recon_size, g.recon_size), dtype=torch.float32, device=proj.device)
    elif tuple(out.shape) != (g.recon_slices, g.recon_size, g.recon_size) or out.dtype != torch.float32:
        raise ValueError("out must be float32 [Nz][N][N]")
    a1 = -1 if a1 is None else int(a1)
    check(lib().vcbct_projector_lengthproj(projector.h, int(ite), C.c_void_p(proj.data_ptr()), int(a0), a1,
                                           C.c_void_p(out.data_ptr()), _stream(stream)))
    return out


def IR_ImageCBCT(before, npf, npf_re, reconScale, wbefore, weight_tv, weight_prior=0.0, out=None, stream=None):
    """One update of the int16 image before [Nz][N][N] (device) with the backprojections npf, npf_re (FP32, same
    shape); returns the FP32 image the reference writes into reconImagefloat."""
    torch = _torch()
    if before.dtype != torch.int16 or before.dim() != 3 or not before.is_contiguous():
        raise ValueError("before must be a contiguous int16 device tensor [Nz][N][N]")
    Nz, N, N2 = before.shape
    if N != N2:
        raise ValueError("the image is N x N per slice")
    for a in (npf, npf_re):
        if a.dtype != torch.float32 or tuple(a.shape) != tuple(before.shape) or not a.is_contiguous():
            raise ValueError("npf and npf_re must be contiguous float32 tensors shaped like before")
    if out is None:
        out = torch.empty(before.shape, dtype=torch.float32, device=before.device)
    check(lib().vcbct_ir_update(N, Nz, float(reconScale), C.c_void_p(before.data_ptr()), C.c_void_p(npf.data_ptr()),
                                C.c_void_p(npf_re.data_ptr()), float(wbefore), float(weight_tv), float(weight_prior),
                                C.c_void_p(out.data_ptr()), _stream(stream)))
    return out


__all__ = ["lengthproj", "IR_ImageCBCT"]
