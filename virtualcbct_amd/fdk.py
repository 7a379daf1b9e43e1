"""FDK: the reference's prepareConstMap / filterSinogram / reconstructImage interface
(F/filterSinogram.h:2-6, F/reconstructImage.h:2-5) over the HIP kernels of libvcbct.so.

The host-array functions take and return numpy arrays laid out like the reference's new3DArray blocks
([th][n][m] projections, [z][y][x] volumes); the `*_device` functions work on torch device tensors.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import P_D, check, lib, ptr
from ._lib import stream_arg as _stream


def _torch():
    import torch
    return torch


def prepareConstMap(projWidth, DETECTOR_PITCH_CM_AT_ISO_CENT=0.032509752):
    """prepareConstMap (F/filterSinogram.cpp:10-65): the W x W FP64 filter matrix h(m - n)."""
    out = np.empty((projWidth, projWidth), np.float64)
    check(lib().vcbct_prepare_const_map(int(projWidth), float(DETECTOR_PITCH_CM_AT_ISO_CENT), ptr(out, P_D)))
    return out


def filter_fft_length(projWidth):
    """The FFT length the filter uses for a panel of projWidth columns (vcbct_filter_fft_length)."""
    return _nonneg(lib().vcbct_filter_fft_length(int(projWidth)))


def _nonneg(rc):
    if rc < 0:
        check(rc)
    return rc


def filter_sinogram_device(proj, xshift, yshift, DETECTOR_PITCH_CM_AT_ISO_CENT=0.032509752, out=None, stream=None):
    """Device form of filterSinogram: proj, out are torch float32 [n][H][W] on the device."""
    torch = _torch()
    n, H, W = proj.shape
    p = proj.contiguous()
    if out is None:
        out = torch.empty_like(p)
    xs = np.ascontiguousarray(xshift, np.float64)
    ys = np.ascontiguousarray(yshift, np.float64)
    if xs.size < n or ys.size < n:
        raise ValueError("xshift/yshift need one entry per projection")
    check(lib().vcbct_filter_sinogram(H, W, n, C.c_void_p(p.data_ptr()), C.c_void_p(out.data_ptr()), ptr(xs, P_D),
                                      ptr(ys, P_D), float(DETECTOR_PITCH_CM_AT_ISO_CENT), _stream(stream)))
    return out


def filterSinogram(projHeight, projWidth, sngrmHeight, projVolume, xshift, yshift,
                   DETECTOR_PITCH_CM_AT_ISO_CENT=0.032509752, device=0):
    """filterSinogram (F/filterSinogram.cpp:67-131) on host arrays: returns fltdVoxData [th][n][m] FP32."""
    torch = _torch()
    pv = np.ascontiguousarray(projVolume, np.float32).reshape(sngrmHeight, projHeight, projWidth)
    d = torch.from_numpy(pv).to(f"cuda:{device}")
    out = filter_sinogram_device(d, xshift, yshift, DETECTOR_PITCH_CM_AT_ISO_CENT)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def reconstruct_image_device(recon, sino, zstart, reconStep, reconScale, xfactor, anglesRad, xshift,
                             DETECTOR_PITCH_CM_AT_ISO_CENT=0.032509752, stream=None):
    """Device form of reconstructImage: accumulates (+=) into recon, a torch float32 [Z][R][R] tensor."""
    Z, R, R2 = recon.shape
    n, H, W = sino.shape
    if R != R2 or not recon.is_contiguous():
        raise ValueError("recon must be a contiguous [Z][R][R] tensor")
    a = np.ascontiguousarray(anglesRad, np.float64)
    xs = np.ascontiguousarray(xshift, np.float64)
    if a.size < n or xs.size < n:
        raise ValueError("angles/xshift need one entry per projection")
    s = sino.contiguous()
    check(lib().vcbct_reconstruct_image(R, Z, int(zstart), float(reconStep), float(reconScale),
                                        C.c_void_p(recon.data_ptr()), H, W, n, C.c_void_p(s.data_ptr()), int(xfactor),
                                        ptr(a, P_D), ptr(xs, P_D), float(DETECTOR_PITCH_CM_AT_ISO_CENT),
                                        _stream(stream)))
    return recon


def reconstructImage(reconSize, reconSlices, zstart, reconStep, reconScale, reconVoxData, projHeight, sngrmWidth,
                     sngrmHeight, sngrmVoxData, xfactor, anglesRad, xshift, DETECTOR_PITCH_CM_AT_ISO_CENT=0.032509752,
                     device=0):
    """reconstructImage (F/reconstructImage.cpp:10-184): accumulates into reconVoxData (numpy FP64 [z][y][x],
    caller-zeroed as in F/main.cpp:168) and returns it."""
    torch = _torch()
    sino = torch.from_numpy(np.ascontiguousarray(sngrmVoxData, np.float32).reshape(sngrmHeight, projHeight,
                                                                                     sngrmWidth)).to(f"cuda:{device}")
    rec = torch.zeros((reconSlices, reconSize, reconSize), dtype=torch.float32, device=f"cuda:{device}")
    reconstruct_image_device(rec, sino, zstart, reconStep, reconScale, xfactor, anglesRad, xshift,
                             DETECTOR_PITCH_CM_AT_ISO_CENT)
    torch.cuda.synchronize()
    reconVoxData += rec.cpu().numpy().astype(np.float64)
    return reconVoxData


def fdk_device(proj, zstart, reconStep, reconScale, reconSize, reconSlices, xfactor, anglesRad, xshift, yshift,
               DETECTOR_PITCH_CM_AT_ISO_CENT=0.032509752, stream=None):
    """filter + backprojection on device tensors; returns the FP32 [Z][R][R] volume (before the int16 cast)."""
    torch = _torch()
    filt = filter_sinogram_device(proj, xshift, yshift, DETECTOR_PITCH_CM_AT_ISO_CENT, stream=stream)
    rec = torch.zeros((reconSlices, reconSize, reconSize), dtype=torch.float32, device=proj.device)
    reconstruct_image_device(rec, filt, zstart, reconStep, reconScale, xfactor, anglesRad, xshift,
                             DETECTOR_PITCH_CM_AT_ISO_CENT, stream=stream)
    return rec


def size_class(xshift):
    """Panel-offset class of the used projections' x shifts (F/projectionDataMVkV.cpp:277-281): "S", "M" or "L"."""
    x = np.ascontiguousarray(xshift, np.float64)
    out = C.create_string_buffer(4)
    check(lib().vcbct_fdk_size_class(len(x), ptr(x, P_D), out))
    return out.value.decode()


def prepareAngles(anglesRad, reconSize, reconScale, SMLClassString, stream=None):
    """The per-voxel angle mask of prepareAngles (F/misc.cpp:87-180): uint8 device tensor [n][R][R]."""
    torch = _torch()
    a = np.ascontiguousarray(anglesRad, np.float64)
    out = torch.empty((len(a), reconSize, reconSize), dtype=torch.uint8, device="cuda")
    check(lib().vcbct_fdk_angle_mask(len(a), ptr(a, P_D), int(reconSize), float(reconScale),
                                     str(SMLClassString).encode(), C.c_void_p(out.data_ptr()), _stream(stream)))
    return out


def half_fan_width(projWidth, xshift):
    """Columns of the padded half-fan panel: W + round(2 max |shift|) (vcbct_fdk_half_fan_width)."""
    xs = np.ascontiguousarray(xshift, np.float64)
    w = lib().vcbct_fdk_half_fan_width(int(projWidth), len(xs), ptr(xs, P_D))
    if w < 0:
        check(w)
    return w


def offset_weight_device(proj, xshift, yshift, taper_px=0.0, DETECTOR_PITCH_CM_AT_ISO_CENT=0.032509752, stream=None):
    """Offset-detector (half-fan) weighting of device projections [n][H][W] before the filter (no reference
    counterpart: the step that the size class and angle mask above are computed for, F/main.cpp:57).  Column
    weights w(t) = sin^2(pi/4 (1 + t/tau)) across the overlap, 0 / 1 outside it, with t the column's distance
    from the central ray (positive towards the long side) and tau = min(taper_px, W/2 - |shift|); conjugate
    rays sum to one.  Acts on the line integral (p ** (2 w)), and pads the short side with columns of p = 1 so
    the panel is symmetric about the central ray.

    Returns (proj_padded [n][H][W + P], xshift_padded, yshift_padded, pitch_iso_padded): filter and
    backproject the padded panel with these (same pixel size, same central ray and row centre)."""
    torch = _torch()
    if proj.dtype != torch.float32 or not proj.is_cuda or proj.dim() != 3 or not proj.is_contiguous():
        raise ValueError("proj must be a contiguous float32 CUDA tensor [n][H][W]")
    n, H, W = proj.shape
    xs = np.ascontiguousarray(xshift, np.float64)
    ys = np.ascontiguousarray(yshift, np.float64)
    if xs.shape != (n,) or ys.shape != (n,):
        raise ValueError(f"xshift and yshift must have {n} entries")
    Wo = half_fan_width(W, xs)
    out = torch.empty((n, H, Wo), dtype=torch.float32, device=proj.device)
    xo, yo = np.empty(n), np.empty(n)
    check(lib().vcbct_fdk_offset_weight(int(H), int(W), int(n), C.c_void_p(proj.data_ptr()), ptr(xs, P_D),
                                        ptr(ys, P_D), float(taper_px), int(Wo), C.c_void_p(out.data_ptr()),
                                        ptr(xo, P_D), ptr(yo, P_D), _stream(stream)))
    return out, xo, yo, DETECTOR_PITCH_CM_AT_ISO_CENT * Wo / W


def fdk_half_fan_device(proj, zstart, reconStep, reconScale, reconSize, reconSlices, xfactor, anglesRad, xshift,
                        yshift, taper_px=0.0, DETECTOR_PITCH_CM_AT_ISO_CENT=0.032509752, stream=None):
    """fdk_device for a half-fan (offset-panel) scan: weighting and padding, then filter + backprojection."""
    p, xs, ys, pit = offset_weight_device(proj, xshift, yshift, taper_px, DETECTOR_PITCH_CM_AT_ISO_CENT, stream)
    return fdk_device(p, zstart, reconStep, reconScale, reconSize, reconSlices, xfactor, anglesRad, xs, ys, pit,
                      stream=stream)


def to_short(recon):
    """(short) truncation toward zero of the FP volume (F/main.cpp:179-185)."""
    r = np.asarray(recon)
    return np.trunc(r).astype(np.int16)


class FdkPlan:
    """Per-geometry FDK plan (tables and filter response built once; execution allocates nothing and can be
    captured in a hipGraph).  The backprojector variant (per-column geometry table or direct, slices per
    thread) is autotuned at creation."""

    def __init__(self, projHeight, projWidth, nProj, reconSize, reconSlices, zstart, reconStep, reconScale, xfactor,
                 anglesRad, xshift, yshift, DETECTOR_PITCH_CM_AT_ISO_CENT=0.032509752, device=0, half_fan=False,
                 taper_px=0.0, window_alpha=0.5):
        """half_fan=True: an offset-panel scan; the plan weights and pads the projections inside its filter call
        (see offset_weight_device) and filters / backprojects the padded panel (width `self.width`).
        window_alpha: the filter's window parameter, 0.5 the reference's kernel (ALPHA, F/physParams.h:39-41),
        1.0 the band-limited ramp without a window (Ram-Lak); not a reference option."""
        a, xs, ys = [np.ascontiguousarray(v, np.float64) for v in (anglesRad, xshift, yshift)]
        if min(a.size, xs.size, ys.size) < nProj:
            raise ValueError("angles/xshift/yshift need one entry per projection")
        h = C.c_void_p()
        args = (int(projHeight), int(projWidth), int(nProj), int(reconSize), int(reconSlices), int(zstart),
                float(reconStep), float(reconScale), int(xfactor), ptr(a, P_D), ptr(xs, P_D), ptr(ys, P_D),
                float(DETECTOR_PITCH_CM_AT_ISO_CENT))
        if half_fan:
            check(lib().vcbct_fdk_plan_create_half_fan(C.byref(h), *args, float(taper_px), int(device)))
        else:
            check(lib().vcbct_fdk_plan_create(C.byref(h), *args, int(device)))
        self.h = h
        if window_alpha != 0.5:
            check(lib().vcbct_fdk_plan_set_window(h, float(window_alpha)))
        self.window_alpha = float(window_alpha)
        self.width = lib().vcbct_fdk_plan_width(h)
        self.shape_proj = (int(nProj), int(projHeight), int(projWidth))
        self.shape_filtered = (int(nProj), int(projHeight), int(self.width))
        self.shape_recon = (int(reconSlices), int(reconSize), int(reconSize))
        self.device = int(device)

    def _check(self, t, shape, name):
        torch = _torch()
        if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous() or tuple(t.shape) != shape:
            raise ValueError(f"{name} must be a contiguous float32 CUDA tensor of shape {shape}")

    @property
    def slices_per_thread(self):
        return lib().vcbct_fdk_plan_slices_per_thread(self.h)

    @property
    def bp_staged(self):
        """True when the backprojector stages each block's projection footprint in LDS (autotuned)."""
        return bool(lib().vcbct_fdk_plan_bp_staged(self.h))

    def filter(self, proj, filtered=None, stream=None):
        """filterSinogram on the device: -log, ramp convolution, cosine weight; returns the filtered set."""
        torch = _torch()
        proj = proj.contiguous()
        self._check(proj, self.shape_proj, "proj")
        if filtered is None:
            filtered = torch.empty(self.shape_filtered, dtype=torch.float32, device=proj.device)
        self._check(filtered, self.shape_filtered, "filtered")
        check(lib().vcbct_fdk_plan_filter(self.h, C.c_void_p(proj.data_ptr()),
                                          C.c_void_p(filtered.data_ptr()), _stream(stream)))
        return filtered

    def backproject(self, filtered, recon, stream=None):
        """reconstructImage on the device: accumulates (+=) into recon (FP32 [Z][R][R])."""
        filtered = filtered.contiguous()
        self._check(filtered, self.shape_filtered, "filtered")
        self._check(recon, self.shape_recon, "recon")
        check(lib().vcbct_fdk_plan_backproject(self.h, C.c_void_p(filtered.data_ptr()),
                                              C.c_void_p(recon.data_ptr()), _stream(stream)))
        return recon

    def execute(self, proj, filtered=None, recon=None, stream=None):
        """filter + backprojection; recon (FP32 [Z][R][R]) is accumulated into (zeroed if created here)."""
        torch = _torch()
        self._check(proj, self.shape_proj, "proj")
        if filtered is None:
            filtered = torch.empty(self.shape_filtered, dtype=torch.float32, device=proj.device)
        if recon is None:
            recon = torch.zeros(self.shape_recon, dtype=torch.float32, device=proj.device)
        self._check(filtered, self.shape_filtered, "filtered")
        self._check(recon, self.shape_recon, "recon")
        check(lib().vcbct_fdk_plan_execute(self.h, C.c_void_p(proj.data_ptr()), C.c_void_p(filtered.data_ptr()),
                                           C.c_void_p(recon.data_ptr()), _stream(stream)))
        return recon

    def close(self):
        if getattr(self, "h", None):
            lib().vcbct_fdk_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
