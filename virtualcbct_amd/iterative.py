"""Iterative-reconstruction pieces of the reference's projector tree, on the device:

  lengthproj    P/lengthproj.cpp:9-131     voxel-driven, chord-weighted backprojection of (65535 - projection)
  IR_ImageCBCT  P/IR_ImageCBCT.cpp:19-94   one MLEM update with a 13-neighbour total-variation penalty

The reference's main program includes both but no longer calls them; they keep its argument meaning
(the projector's geometry for lengthproj, a [Nz][N][N] int16 image for the update) and give its results
bit for bit.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib
from ._lib import stream_arg as _stream


def _torch():
    import torch
    return torch


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


# Water by mass (the reference's water phantom: H 0.111, O 0.889; material order H, C, N, O, P, Ca).
WATER_FRACTIONS = (0.111, 0.0, 0.0, 0.889, 0.0, 0.0)


class MlemTV:
    """Iterative reconstruction on the projector's own kernels: the reference's MLEM update with its
    13-neighbour TV penalty (IR_ImageCBCT, P/IR_ImageCBCT.cpp:19-94, bit-exact on the device) applied to
    backprojections formed by the ray-driven adjoint of the forward projector, not by lengthproj's voxel-driven
    chords.  With A the monochromatic log projection of the image x (int16, density x 1000 of one material mix:
    six maps f_m x / 1000; p = sum_m mu_m L_m / xf, linear in x) and A^T its exact transpose (the adjoint
    scatter), one iteration is
        npf = A^T (A x_k),   x_{k+1} = IR_ImageCBCT(x_k, npf, npf_re, ...),   npf_re = A^T b  (once),
    the one-step-late MLEM-TV of the reference's update, x (wbefore + (1 - wbefore) npf_re / (npf + lambda)),
    lambda = weight_tv TV'(x) + 1e-6, clamped and masked as the reference does; the image is truncated to int16
    between iterations, as the reference's short image.  b: measured projections as the log projection of the
    same operator ([n_angles][H][W] FP32).  There is no reference loop around the two pieces (P/main.cpp no longer
    calls them); this one is built on the forward and adjoint kernels (A x: one traversal keeping its line
    integrals; A^T r: one scatter, in a single channel since the monochromatic table has rank one)."""

    def __init__(self, projector, mu, b, fractions=WATER_FRACTIONS, xfactor=1.0, wbefore=0.0, weight_tv=0.0):
        torch = _torch()
        g = projector.g
        self.P = projector
        self.shape = (g.recon_slices, g.recon_size, g.recon_size)
        mu = np.asarray(mu, np.float64).reshape(g.num_material, 1)
        projector.set_spectrum(mu.astype(np.float32), np.ones(1, np.float32))
        f = torch.tensor(np.asarray(fractions, np.float64)[: g.num_material], dtype=torch.float32,
                         device=f"cuda:{projector.device}")
        self.f = f.view(-1, 1, 1, 1)
        self.xf = float(xfactor)
        self.wbefore, self.weight_tv, self.scale = float(wbefore), float(weight_tv), float(g.recon_scale)
        self.maps = torch.empty((g.num_material,) + self.shape, dtype=torch.float32, device=f"cuda:{projector.device}")
        # A^T reads the line integrals A kept (with one energy bin the coefficients are mu_m / xf whatever they
        # are); a first projection of the empty image provides them
        self.A(torch.zeros(self.shape, dtype=torch.int16, device=self.maps.device))
        self.npf_re = self.AT(b)

    def A(self, x):
        """The log projection of the int16 (or FP32) image x: maps f_m x / 1000, then p = sum_m mu_m L_m / xf."""
        torch = _torch()
        torch.mul(self.f, x.to(torch.float32).unsqueeze(0) * 1e-3, out=self.maps)
        self.P.set_volume(self.maps)
        p, self._L = self.P.logproj_save(self.xf)
        return p

    def AT(self, r):
        """The transpose: sum_m f_m (A_m^T r) / 1000, an [Nz][N][N] FP32 image."""
        g = self.P.logproj_backward_saved(self._L, r.contiguous(), self.xf)
        return ((g * self.f).sum(0) * 1e-3).contiguous()

    def step(self, x):
        """One iteration from the int16 image x; returns the updated FP32 image (IR_ImageCBCT's output)."""
        npf = self.AT(self.A(x))
        return IR_ImageCBCT(x, npf, self.npf_re, self.scale, self.wbefore, self.weight_tv)

    def run(self, x0, iters):
        """iters iterations from the int16 image x0; returns (the last FP32 image, the int16 image)."""
        torch = _torch()
        x = x0.to(torch.int16).contiguous()
        out = None
        for _ in range(iters):
            out = self.step(x)
            x = out.to(torch.int16)  # the reference's short image (truncation towards zero)
        return out, x


__all__ = ["lengthproj", "IR_ImageCBCT", "MlemTV", "WATER_FRACTIONS"]
