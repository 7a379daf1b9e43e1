"""Organ-label phantom -> six elemental maps (H, C, N, O, P, Ca), the step before the projector.

The reference does this offline on the host (ICRPphantoms_2mmslice/To_material_distribution_for512_random.cpp):
it perturbs the ICRP media compositions (:117-165), then scans every organ row for every voxel
(:212-226) and writes per-slice material files (:242-274).  Here the tables become a 256-entry lookup
table on the host (vcbct_material_table) and the voxel work runs on the device, either in the
reference's per-slice layout (labels_to_maps) or straight into a Projector's volume
(Projector.set_volume_labels), with the crop / pad from the phantom's slice size to the projector's
N x N fused in.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._lib import P_F, P_U16, VcbctError, check, lib
from ._lib import stream_arg as _stream

MATERIALS = ("H", "C", "N", "O", "P", "Ca")


def _torch():
    import torch
    return torch


def material_table(media_csv, organs_csv, randomize=True, seed=1):
    """(lut, media_weights): lut uint16 [256][8], the six map values (g/cc x 1000) of every 8-bit organ
    label (columns 6-7 zero); media_weights float32 [n_media][6], the H C N O P Ca weight % used.

    randomize=True redraws the weights as the reference does (R/ :117-150) from glibc's rand() sequence
    after srand(seed); seed=1 reproduces the reference's output."""
    lut = np.zeros((256, 8), np.uint16)
    w = np.zeros((4096, 6), np.float32)
    n = lib().vcbct_material_table(os.fspath(media_csv).encode(), os.fspath(organs_csv).encode(),
                                   int(bool(randomize)), C.c_uint(int(seed)), lut.ctypes.data_as(P_U16),
                                   w.ctypes.data_as(P_F), len(w))
    if n < 0:
        check(n)
    return lut, w[:n].copy()


_LUTS = {}


def device_lut(lut, device=0):
    """The [256][8] table as a contiguous int16 device tensor (bit pattern of the uint16 values).  Tables
    made from host arrays are kept (a few, per device) so that a launch on any stream can still read them."""
    torch = _torch()
    if isinstance(lut, np.ndarray):
        a = np.ascontiguousarray(lut, np.uint16)
        if a.size != 2048:
            raise ValueError("the label table has 256 x 8 entries")
        key = (int(device), a.tobytes())
        t = _LUTS.get(key)
        if t is None:
            if len(_LUTS) >= 16:
                _LUTS.pop(next(iter(_LUTS)))
            t = torch.from_numpy(a.view(np.int16).reshape(-1).copy()).to(f"cuda:{device}")
            _LUTS[key] = t
        return t
    if lut.numel() != 2048 or lut.element_size() != 2 or not lut.is_contiguous():
        raise ValueError("the label table has 256 x 8 16-bit entries, contiguous")
    return lut


def _check_labels(labels):
    torch = _torch()
    if labels.dtype != torch.uint8 or labels.dim() != 3 or not labels.is_contiguous() or not labels.is_cuda:
        raise ValueError("labels must be a contiguous uint8 device tensor [nz][nrow][ncol]")
    return labels.shape


def labels_to_maps(labels, lut, out=None, stream=None):
    """labels uint8 [nz][nrow][ncol] (device) -> int16 [nz][6][nrow][ncol] device tensor holding the uint16
    map values; slice iz is the reference's material file <tag>_MD_<ncol>_<nrow>_<iz>.raw."""
    torch = _torch()
    nz, nrow, ncol = _check_labels(labels)
    dl = device_lut(lut, labels.device.index or 0)
    if out is None:
        out = torch.empty((nz, 6, nrow, ncol), dtype=torch.int16, device=labels.device)
    elif tuple(out.shape) != (nz, 6, nrow, ncol) or out.element_size() != 2 or not out.is_contiguous():
        raise ValueError("out must be a contiguous 16-bit tensor [nz][6][nrow][ncol]")
    check(lib().vcbct_labels_to_maps(C.c_void_p(labels.data_ptr()), ncol, nrow, nz, C.c_void_p(dl.data_ptr()),
                                     C.c_void_p(out.data_ptr()), _stream(stream)))
    return out


def centred_placement(ncol, nrow, N):
    """(col0, row0) that centre an ncol x nrow slice on an N x N grid (crop where larger, pad where smaller)."""
    return (ncol - N) // 2, (nrow - N) // 2


def write_md_slices(maps, dirpath, tag="AM"):
    """Per-slice material files as the reference names them (R/ :244-245): <dir>/<tag>_MD_<ncol>_<nrow>_<iz>.raw,
    each [6][nrow][ncol] int16.  maps: host or device [nz][6][nrow][ncol]."""
    a = maps.cpu().numpy() if hasattr(maps, "cpu") else np.asarray(maps)
    nz, _, nrow, ncol = a.shape
    os.makedirs(dirpath, exist_ok=True)
    paths = []
    for iz in range(nz):
        p = os.path.join(dirpath, f"{tag}_MD_{ncol}_{nrow}_{iz:03d}.raw")
        np.ascontiguousarray(a[iz]).view(np.int16).tofile(p)
        paths.append(p)
    return paths


__all__ = ["MATERIALS", "material_table", "device_lut", "labels_to_maps", "centred_placement", "write_md_slices",
           "VcbctError"]
