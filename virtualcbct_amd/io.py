"""Reference file formats over the C ABI: the projector and FDK input txt files, flexmaps, elemental slice
stacks and the spectrum / xraydb tables (attenuation_coefficient)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import P_D, P_F, P_I, P_U16, check, lib, ptr


def _s(b):
    return bytes(b).split(b"\0")[0].decode()


def read_proj_config(path):
    """loadData of the projector (P/projectionData.cpp:63-102) -> dict of the 15 fields + physParams."""
    c = _lib.vcbct_proj_config()
    check(lib().vcbct_read_proj_config(path.encode(), C.byref(c)))
    d = {k: getattr(c, k) for k, _ in c._fields_}
    for k in ("spectrum", "flexmap", "mat_prefix", "outputname", "mat_type"):
        d[k] = _s(getattr(c, k))
    d["_struct"] = c
    return d


def read_flexmap(path, capacity=100000):
    a, x, y = np.empty(capacity), np.empty(capacity), np.empty(capacity)
    n = lib().vcbct_read_flexmap(path.encode(), ptr(a, P_D), ptr(x, P_D), ptr(y, P_D), capacity)
    if n < 0:
        check(n)
    return a[:n].copy(), x[:n].copy(), y[:n].copy()


def slice_window(cfg):
    lo, hi = C.c_int(), C.c_int()
    check(lib().vcbct_proj_slice_window(C.byref(cfg["_struct"]), C.byref(lo), C.byref(hi)))
    return lo.value, hi.value


def load_material_slices(prefix, linemin, linemax, recon_size, num_material, missing_policy=0):
    vol = np.empty((linemax - linemin, num_material, recon_size, recon_size), np.uint16)
    nm = C.c_int(0)
    check(lib().vcbct_load_material_slices(prefix.encode(), linemin, linemax, recon_size, num_material,
                                           missing_policy, ptr(vol, P_U16), C.byref(nm)))
    return vol, nm.value


def read_fdk_config(path):
    """loadData of cbct2 (F/projectionDataMVkV.cpp:51-97)."""
    c = _lib.vcbct_fdk_config()
    check(lib().vcbct_read_fdk_config(path.encode(), C.byref(c)))
    d = {k: getattr(c, k) for k, _ in c._fields_}
    for k in ("kvmv", "proj_path", "ext", "flexmap", "tgt_class", "output"):
        d[k] = _s(getattr(c, k))
    return d


def read_fdk_angles(path, total, tgt_class="*"):
    a, x, y = np.empty(total), np.empty(total), np.empty(total)
    n = lib().vcbct_read_fdk_angles(path.encode(), tgt_class.encode(), int(total), ptr(a, P_D), ptr(x, P_D),
                                    ptr(y, P_D))
    if n < 0:
        check(n)
    return a[:n].copy(), x[:n].copy(), y[:n].copy()


def attenuation_coefficient(spectrum_csv, xraydb_dir, num_material=6, nthin=1):
    """attenuation_coefficient (P/prior_weight_production.cpp:200-523): returns (Attenuation [mat][NE], pdf [NE])."""
    at = np.zeros(1024 * num_material, np.float32)
    pdf = np.zeros(1024, np.float32)
    ne = lib().vcbct_attenuation_table(spectrum_csv.encode(), xraydb_dir.encode(), num_material, nthin,
                                       ptr(at, P_F), ptr(pdf, P_F), 1024)
    if ne < 0:
        check(ne)
    return at[: num_material * ne].reshape(num_material, ne).copy(), pdf[:ne].copy()


def trig_table(W, H, angles_rad, xshift, yshift, det_pitch_cm=0.05, det_pixel_num=800.0):
    """prepareTrigonometricTable (P/misc.cpp:9-37) -> cos, sin, XcenterShift, YcenterShift."""
    n = len(angles_rad)
    a, x, y = [np.ascontiguousarray(v, np.float64) for v in (angles_rad, xshift, yshift)]
    outs = [np.empty(n) for _ in range(4)]
    check(lib().vcbct_trig_table(int(W), int(H), n, ptr(a, P_D), ptr(x, P_D), ptr(y, P_D), float(det_pitch_cm),
                                 float(det_pixel_num), *[ptr(o, P_D) for o in outs]))
    return outs
