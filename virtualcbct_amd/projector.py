"""Forward projector: the reference's initializeGPU / reprojectionOnGPU / terminateGPU interface
(P/reprojectionOnGPU.h:2-19) over the HIP kernels of libvcbct.so.

Device tensors are torch tensors on the ROCm device (PyTorch is the allocator and stream provider);
the computation is the hand-written HIP code behind the C ABI.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import P_D, P_F, P_U16, check, lib, ptr
from ._lib import stream_arg as _stream


def _torch():
    import torch
    return torch


class Projector:
    """Device state of initializeGPU (P/reprojectionOnGPU.cu:720-825).

    Arguments follow initializeGPU's (P/reprojectionOnGPU.h:2-8); `device` replaces the reference's
    hard-coded device 1.
    """

    def __init__(self, reconSize, reconScale, reconSlices, reconStep, projImgHeight, projImgWidth,
                 totalProjNumber, cosTable, sinTable, XcenterShift, YcenterShift, x_reg=0.0, y_reg=0.0,
                 z_reg=0.0, DETECTOR_PITCH_CM=0.05, DETECTOR_PITCH_CM_AT_ISO_CENT=0.032509752,
                 DIST_BTWN_SRC_AND_ISOCENT_CM=100.0, DETECTOR_PIXEL_NUM=800.0, num_material=6, device=0):
        self.g = _lib.vcbct_geometry(int(reconSize), float(reconScale), int(reconSlices), float(reconStep),
                                     int(projImgHeight), int(projImgWidth), int(totalProjNumber), float(x_reg),
                                     float(y_reg), float(z_reg), float(DETECTOR_PITCH_CM),
                                     float(DETECTOR_PITCH_CM_AT_ISO_CENT), float(DIST_BTWN_SRC_AND_ISOCENT_CM),
                                     float(DETECTOR_PIXEL_NUM), int(num_material))
        tabs = [np.ascontiguousarray(t, np.float64) for t in (cosTable, sinTable, XcenterShift, YcenterShift)]
        for t in tabs:
            if t.shape != (int(totalProjNumber),):
                raise ValueError("angle tables must have totalProjNumber entries")
        self.device = int(device)
        h = C.c_void_p()
        check(lib().vcbct_projector_create(C.byref(h), C.byref(self.g), *[ptr(t, P_D) for t in tabs], self.device))
        self.h = h
        self.NE = 0
        self.spectrum_generation = 0
        self._vol_keep = None

    # shape helpers
    @property
    def shape_proj(self):
        return (self.g.n_angles, self.g.proj_height, self.g.proj_width)

    @property
    def shape_maps(self):
        return (self.g.num_material, self.g.recon_slices, self.g.recon_size, self.g.recon_size)

    def close(self):
        if getattr(self, "h", None):
            lib().vcbct_projector_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------ uploads
    def set_volume(self, vol, stream=None):
        """vol: uint16 [z][mat][y][x] (numpy host array or torch int16/uint16 device tensor, the reference
        layout), or a torch float32 device tensor [mat][z][y][x] (elemental maps of the differentiable
        path)."""
        n = self.g
        if isinstance(vol, np.ndarray):
            v = np.ascontiguousarray(vol, np.uint16)
            if v.size != n.recon_slices * n.num_material * n.recon_size ** 2:
                raise ValueError(f"volume has {v.size} voxels, geometry needs "
                                 f"{n.recon_slices}x{n.num_material}x{n.recon_size}^2")
            check(lib().vcbct_projector_set_volume_u16(self.h, v.ctypes.data_as(C.c_void_p), 0, _stream(stream)))
            return
        torch = _torch()
        if vol.dtype == torch.float32:
            if tuple(vol.shape) != self.shape_maps or not vol.is_contiguous():
                raise ValueError(f"FP32 maps must be contiguous {self.shape_maps}")
            check(lib().vcbct_projector_set_volume_f32(self.h, C.c_void_p(vol.data_ptr()), _stream(stream)))
            self._vol_keep = vol
            return
        if vol.dtype not in (torch.int16, torch.uint16) or not vol.is_contiguous():
            raise ValueError("device volume must be a contiguous int16/uint16 tensor")
        if vol.numel() != n.recon_slices * n.num_material * n.recon_size ** 2:
            raise ValueError("device volume size does not match the geometry")
        check(lib().vcbct_projector_set_volume_u16(self.h, C.c_void_p(vol.data_ptr()), 1, _stream(stream)))

    def set_volume_labels(self, labels, lut, col0=None, row0=None, slice0=0, stream=None):
        """Organ-label phantom straight into the device volume: labels uint8 [nz][nrow][ncol] (device), lut
        [256][8] as material_maps.material_table returns it.  Voxel (x, y, z) takes the label at column
        x + col0, row y + row0, slice z + slice0 (label 0 outside the phantom); col0/row0 default to the
        centring offsets."""
        from . import material_maps
        nz, nrow, ncol = material_maps._check_labels(labels)
        c0, r0 = material_maps.centred_placement(ncol, nrow, self.g.recon_size)
        col0 = c0 if col0 is None else int(col0)
        row0 = r0 if row0 is None else int(row0)
        dl = material_maps.device_lut(lut, self.device)
        check(lib().vcbct_projector_set_volume_labels(self.h, C.c_void_p(labels.data_ptr()), ncol, nrow, nz, col0,
                                                      row0, int(slice0), C.c_void_p(dl.data_ptr()),
                                                      _stream(stream)))

    def get_volume(self, stream=None):
        """The u16 device volume in the reference layout [z][mat][y][x] (int16 device tensor of the uint16
        values): the projector-ready slices, e.g. of a label phantom."""
        torch = _torch()
        n = self.g
        out = torch.empty((n.recon_slices, n.num_material, n.recon_size, n.recon_size), dtype=torch.int16,
                          device=f"cuda:{self.device}")
        check(lib().vcbct_projector_get_volume_u16(self.h, C.c_void_p(out.data_ptr()), _stream(stream)))
        return out

    def set_spectrum(self, Attenuation, pdf_pe):
        """Attenuation [mat][NE] (or flat mat*NE) cm^2/g and pdf_pe [NE], as attenuation_coefficient writes."""
        pdf = np.ascontiguousarray(pdf_pe, np.float32).ravel()
        ne = pdf.size
        at = np.ascontiguousarray(Attenuation, np.float32).ravel()
        if at.size < self.g.num_material * ne:
            raise ValueError("attenuation table smaller than num_material x NE")
        at = np.ascontiguousarray(at[: self.g.num_material * ne])
        check(lib().vcbct_projector_set_spectrum(self.h, ptr(at, P_F), ptr(pdf, P_F), ne))
        self.NE = ne
        self.spectrum = (at.reshape(self.g.num_material, ne).copy(), pdf.copy())
        # saved autograd state refers to the spectrum it was computed with; a backward pass checks this
        self.spectrum_generation += 1

    def set_spectra(self, spectra):
        """Spectra of the multi-spectrum loss: a list of (Attenuation [mat][NE_s], pdf_pe [NE_s]) pairs (1 to 4),
        independent of set_spectrum's single spectrum."""
        nm = self.g.num_material
        ats, pds, nes = [], [], []
        for atten, pdf_pe in spectra:
            pdf = np.ascontiguousarray(pdf_pe, np.float32).ravel()
            at = np.ascontiguousarray(atten, np.float32).ravel()
            if at.size < nm * pdf.size:
                raise ValueError("attenuation table smaller than num_material x NE")
            ats.append(at[: nm * pdf.size])
            pds.append(pdf)
            nes.append(pdf.size)
        at = np.ascontiguousarray(np.concatenate(ats))
        pd = np.ascontiguousarray(np.concatenate(pds))
        ne = np.ascontiguousarray(nes, np.int32)
        check(lib().vcbct_projector_set_spectra(self.h, len(nes), ptr(at, P_F), ptr(pd, P_F),
                                                ne.ctypes.data_as(_lib.P_I)))
        self.nspec = len(nes)

    @property
    def multi_channels(self):
        rc = lib().vcbct_projector_multi_channels(self.h)
        if rc < 0:
            check(rc)
        return rc

    def multi_loss_grad(self, targets, xfactor, grad=None, loss=None, stream=None):
        """loss = sum over the set_spectra spectra and the rays of (p_s - t_s)^2 for the FP32 maps set with
        set_volume, and its gradient [mat][z][y][x]: one traversal and one scatter for all spectra.  targets:
        contiguous FP32 [nspec][n_angles][H][W] on the device.  Returns (grad, loss) (loss: one FP64 value)."""
        torch = _torch()
        nspec = getattr(self, "nspec", 0)
        self._check_dev(targets, (nspec,) + self.shape_proj, "targets")
        if grad is None:
            grad = torch.empty(self.shape_maps, dtype=torch.float32, device=f"cuda:{self.device}")
        self._check_dev(grad, self.shape_maps, "grad")
        if loss is None:
            loss = torch.zeros(1, dtype=torch.float64, device=f"cuda:{self.device}")
        self._check_dev(loss, (1,), "loss", torch.float64)
        check(lib().vcbct_projector_multi_loss_grad(self.h, C.c_void_p(targets.data_ptr()),
                                                    C.c_void_p(grad.data_ptr()), C.c_void_p(loss.data_ptr()),
                                                    float(xfactor), _stream(stream)))
        return grad, loss

    @property
    def forward_channels(self):
        """Spectral channels the forward traverses in place of the maps (2 or 4 FP32 per voxel), 0 for the maps."""
        rc = lib().vcbct_projector_forward_channels(self.h)
        if rc < 0:
            check(rc)
        return rc

    @property
    def adjoint_channels(self):
        """Channels the adjoint scatters: num_material, or the rank of the attenuation table's spectral basis."""
        rc = lib().vcbct_projector_adjoint_channels(self.h)
        if rc < 0:
            check(rc)
        return rc

    # ------------------------------------------------------------------ compute
    def forward(self, xfactor, out=None, a0=0, a1=None, stream=None):
        """Polychromatic intensity projections [n_angles][H][W] (torch float32 on the device)."""
        torch = _torch()
        if out is None:
            out = torch.empty(self.shape_proj, dtype=torch.float32, device=f"cuda:{self.device}")
        a1 = self.g.n_angles if a1 is None else a1
        check(lib().vcbct_projector_forward(self.h, C.c_void_p(out.data_ptr()), float(xfactor), int(a0), int(a1),
                                            _stream(stream)))
        return out

    def line_integrals(self, stream=None):
        torch = _torch()
        L = torch.empty(self.shape_proj + (self.g.num_material,), dtype=torch.float32, device=f"cuda:{self.device}")
        check(lib().vcbct_projector_line_integrals(self.h, C.c_void_p(L.data_ptr()), _stream(stream)))
        return L

    def segment_counts(self, stream=None):
        """Charged segments per ray (FP32 [n_angles][H][W]); independent of the volume's values."""
        torch = _torch()
        c = torch.empty(self.shape_proj, dtype=torch.float32, device=f"cuda:{self.device}")
        check(lib().vcbct_projector_segment_counts(self.h, C.c_void_p(c.data_ptr()), _stream(stream)))
        return c

    def set_tile(self, row_groups=0, cols=0, order=0, split=0):
        """Fixes the tile shape of the intensity / log-projection launches (row_groups x cols waves per
        workgroup, 2, 4 or 8; order 0: angles inside row blocks, 1: row blocks inside angles; split 0: columns,
        1: angles divided among the XCDs).  row_groups=0 hands it back to the autotuner."""
        check(lib().vcbct_projector_set_tile(self.h, int(row_groups), int(cols), int(order), int(split)))

    def set_adjoint_tile(self, row_groups=0, cols=0, order=0, split=0):
        """Fixes the tile shape of the adjoint's traversal kernels (row_groups x cols = 4 waves per workgroup);
        row_groups=0 hands it back to the autotuner."""
        check(lib().vcbct_projector_set_adjoint_tile(self.h, int(row_groups), int(cols), int(order), int(split)))

    @property
    def tile(self):
        """(row_groups, cols, order, split) of the last intensity / log-projection launch."""
        a = (C.c_int * 4)()
        check(lib().vcbct_projector_get_tile(self.h, a))
        return tuple(a)

    def logproj(self, xfactor, out=None, stream=None):
        torch = _torch()
        if out is None:
            out = torch.empty(self.shape_proj, dtype=torch.float32, device=f"cuda:{self.device}")
        check(lib().vcbct_projector_logproj(self.h, C.c_void_p(out.data_ptr()), float(xfactor), _stream(stream)))
        return out

    def adjoint(self, target, xfactor, stream=None):
        """Gradient of sum (p - target)^2 w.r.t. the FP32 maps set with set_volume; returns (grad, loss)."""
        torch = _torch()
        t = target.contiguous().to(torch.float32)
        grad = torch.empty(self.shape_maps, dtype=torch.float32, device=f"cuda:{self.device}")
        loss = torch.zeros(1, dtype=torch.float64, device=f"cuda:{self.device}")
        check(lib().vcbct_projector_adjoint(self.h, C.c_void_p(t.data_ptr()), C.c_void_p(grad.data_ptr()),
                                            C.c_void_p(loss.data_ptr()), float(xfactor), _stream(stream)))
        return grad, loss

    @property
    def n_rays(self):
        return self.g.n_angles * self.g.proj_height * self.g.proj_width

    def _check_dev(self, t, shape, name, dtype=None):
        """t must be a contiguous tensor of `shape` (and dtype, FP32 by default) on this projector's device:
        the kernels index it by the projector's geometry, not by the tensor's size."""
        torch = _torch()
        dtype = torch.float32 if dtype is None else dtype
        if (t.dtype != dtype or not t.is_cuda or t.device.index != self.device or not t.is_contiguous()
                or tuple(t.shape) != tuple(shape)):
            raise ValueError(f"{name} must be a contiguous {dtype} tensor of shape {tuple(shape)} on cuda:{self.device}")

    def logproj_save(self, xfactor, out=None, stream=None):
        """(log projection, its six line integrals per ray [nmat][n_angles*H*W]) for logproj_backward_saved."""
        torch = _torch()
        if out is None:
            out = torch.empty(self.shape_proj, dtype=torch.float32, device=f"cuda:{self.device}")
        self._check_dev(out, self.shape_proj, "out")
        L = torch.empty((self.g.num_material, self.n_rays), dtype=torch.float32, device=f"cuda:{self.device}")
        check(lib().vcbct_projector_logproj_save(self.h, C.c_void_p(out.data_ptr()), C.c_void_p(L.data_ptr()),
                                                 float(xfactor), _stream(stream)))
        return out, L

    def logproj_backward_saved(self, L, grad_p, xfactor, stream=None):
        """The backward pass from the line integrals kept by logproj_save (the volume is not read)."""
        torch = _torch()
        if not torch.is_tensor(grad_p) or grad_p.numel() != self.n_rays or not grad_p.is_cuda \
                or grad_p.device.index != self.device or not grad_p.dtype.is_floating_point:
            raise ValueError(f"grad_p must be a floating-point tensor of {self.n_rays} rays on cuda:{self.device}")
        gp = grad_p.contiguous().to(torch.float32)
        self._check_dev(L, (self.g.num_material, self.n_rays), "L (the line integrals returned by logproj_save)")
        grad = torch.empty(self.shape_maps, dtype=torch.float32, device=f"cuda:{self.device}")
        check(lib().vcbct_projector_logproj_backward_saved(self.h, C.c_void_p(L.data_ptr()),
                                                           C.c_void_p(gp.data_ptr()), C.c_void_p(grad.data_ptr()),
                                                           float(xfactor), _stream(stream)))
        return grad

    def logproj_backward(self, grad_p, xfactor, stream=None):
        torch = _torch()
        gp = grad_p.contiguous().to(torch.float32)
        grad = torch.empty(self.shape_maps, dtype=torch.float32, device=f"cuda:{self.device}")
        check(lib().vcbct_projector_logproj_backward(self.h, C.c_void_p(gp.data_ptr()), C.c_void_p(grad.data_ptr()),
                                                     float(xfactor), _stream(stream)))
        return grad


# ---------------------------------------------------------------------------------------------------
# The reference's module-level interface (P/reprojectionOnGPU.h:2-19), one projector per process.

_current: Projector | None = None


def initializeGPU(reconSize, reconScale, reconSlices, reconStep, projImgHeight, projImgWidth, totalProjNumber,
                  cosTable, sinTable, XcenterShift, YcenterShift, x_reg, y_reg, z_reg, DETECTOR_PITCH_CM,
                  DETECTOR_PITCH_CM_AT_ISO_CENT, DIST_BTWN_SRC_AND_ISOCENT_CM, DETECTOR_PIXEL_NUM, num_material,
                  device=0):
    """initializeGPU (P/reprojectionOnGPU.cu:720-825)."""
    global _current
    if _current is not None:
        _current.close()
    _current = Projector(reconSize, reconScale, reconSlices, reconStep, projImgHeight, projImgWidth,
                         totalProjNumber, cosTable, sinTable, XcenterShift, YcenterShift, x_reg, y_reg, z_reg,
                         DETECTOR_PITCH_CM, DETECTOR_PITCH_CM_AT_ISO_CENT, DIST_BTWN_SRC_AND_ISOCENT_CM,
                         DETECTOR_PIXEL_NUM, num_material, device)
    return _current


def terminateGPU():
    """terminateGPU (P/reprojectionOnGPU.cu:827-834)."""
    global _current
    if _current is not None:
        _current.close()
        _current = None


def reprojectionOnGPU(reconImageshort, xfactor, Attenuation, pdf_pe, NE, outputname=None, endProjNumber=None):
    """reprojectionOnGPU (P/reprojectionOnGPU.cu:838-1100): projects the uint16 [z][mat][y][x] volume with the
    spectrum and, like the reference, writes `<outputname>.raw` (FP32 [angle][row][col]) when a name is
    given.  Returns the projections as a host numpy array."""
    if _current is None:
        raise RuntimeError("initializeGPU was not called")
    p = _current
    vol = np.ascontiguousarray(reconImageshort, np.uint16)
    at = np.ascontiguousarray(np.asarray(Attenuation, np.float32).ravel()[: p.g.num_material * NE])
    pdf = np.ascontiguousarray(np.asarray(pdf_pe, np.float32).ravel()[:NE])
    # the reference computes and writes angles [0, endProjNumber) (P/reprojectionOnGPU.cu:976)
    n_out = p.g.n_angles if endProjNumber is None else int(endProjNumber)
    out = np.empty((n_out,) + p.shape_proj[1:], np.float32)
    check(lib().vcbct_reprojection_angles(p.h, ptr(vol, P_U16), ptr(at, P_F), ptr(pdf, P_F), int(NE),
                                          float(xfactor), n_out, ptr(out, P_F)))
    if outputname:
        out.tofile(f"{outputname}.raw")
    return out
