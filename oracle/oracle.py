"""ctypes bindings of the FP64 CPU oracle (oracle/liboracle.so) and of the reference's own CPU code
built from its sources (oracle/_ref/*.so).

TEST INFRASTRUCTURE ONLY.  Only the tests, the smoke check and the benchmark's CPU-baseline leg import
this module, and only as the checker; the product package never does.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_U16 = C.POINTER(C.c_uint16)
_I32 = C.POINTER(C.c_int32)
_CH = C.POINTER(C.c_char)


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


class orc_geom(C.Structure):
    _fields_ = [("N", C.c_int), ("s", C.c_double), ("Nz", C.c_int), ("t", C.c_double),
                ("H", C.c_int), ("W", C.c_int), ("n_ang", C.c_int),
                ("x_reg", C.c_double), ("y_reg", C.c_double), ("z_reg", C.c_double),
                ("det_pitch", C.c_double), ("det_pitch_iso", C.c_double), ("sad", C.c_double),
                ("det_pixel_num", C.c_double), ("nmat", C.c_int)]


@dataclass
class Geometry:
    """The initializeGPU argument block (P/reprojectionOnGPU.h:2-8) plus the trig tables."""
    N: int
    s: float
    Nz: int
    t: float
    H: int
    W: int
    cos_t: np.ndarray
    sin_t: np.ndarray
    xc: np.ndarray
    yc: np.ndarray
    x_reg: float = 0.0
    y_reg: float = 0.0
    z_reg: float = 0.0
    det_pitch: float = 0.05
    det_pitch_iso: float = 0.032509752
    sad: float = 100.0
    det_pixel_num: float = 800.0
    nmat: int = 6

    @property
    def n_ang(self):
        return len(self.cos_t)

    def c(self):
        return orc_geom(self.N, self.s, self.Nz, self.t, self.H, self.W, self.n_ang, self.x_reg,
                        self.y_reg, self.z_reg, self.det_pitch, self.det_pitch_iso, self.sad,
                        self.det_pixel_num, self.nmat)

    def tables(self):
        return [np.ascontiguousarray(a, dtype=np.float64) for a in (self.cos_t, self.sin_t, self.xc, self.yc)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run `make -C oracle`)")
        L = C.CDLL(path)
        L.orc_attenuation_table.restype = C.c_int
        _lib = L
    return _lib


def trig_table_proj(W, H, angles_rad, xshift, yshift, det_pitch=0.05, det_pixel_num=800.0):
    n = len(angles_rad)
    a = np.ascontiguousarray(angles_rad, np.float64)
    xs = np.ascontiguousarray(xshift, np.float64)
    ys = np.ascontiguousarray(yshift, np.float64)
    out = [np.zeros(n) for _ in range(4)]
    lib().orc_trig_table_proj(W, H, n, _p(a, _D), _p(xs, _D), _p(ys, _D), C.c_double(det_pitch),
                              C.c_double(det_pixel_num), *[_p(o, _D) for o in out])
    return out


def attenuation_table(spectrum_csv, xraydb_dir, nmat=6, nthin=1):
    atten = np.zeros(1024 * nmat, np.float32)
    pdf = np.zeros(1024, np.float32)
    ne = lib().orc_attenuation_table(spectrum_csv.encode(), xraydb_dir.encode(), nmat, nthin,
                                     _p(atten, _F), _p(pdf, _F), 1024)
    if ne <= 0:
        raise RuntimeError(f"orc_attenuation_table failed: {ne}")
    return atten[: nmat * ne].reshape(nmat, ne).copy(), pdf[:ne].copy()


def forward_compat(g: Geometry, vol_u16, atten, pdf, xfactor, a0=0, a1=None):
    a1 = g.n_ang if a1 is None else a1
    vol = np.ascontiguousarray(vol_u16, np.uint16)
    at = np.ascontiguousarray(atten, np.float32)
    pd = np.ascontiguousarray(pdf, np.float32)
    out = np.zeros((g.n_ang, g.H, g.W), np.float32)
    cs, sn, xc, yc = g.tables()
    gc = g.c()
    lib().orc_forward_compat(C.byref(gc), _p(cs, _D), _p(sn, _D), _p(xc, _D), _p(yc, _D), _p(vol, _U16),
                             _p(at, _F), _p(pd, _F), len(pd), C.c_double(xfactor), a0, a1, _p(out, _F))
    return out


def line_integrals(g: Geometry, vol, a0=0, a1=None):
    """Returns (L [ang,H,W,nmat] FP64, nseg [ang,H,W], leng [ang,H,W]) for angles [a0,a1) (indexed from 0)."""
    a1 = g.n_ang if a1 is None else a1
    cs, sn, xc, yc = [t[a0:a1].copy() for t in g.tables()]
    na = a1 - a0
    L = np.zeros((na, g.H, g.W, g.nmat))
    ns = np.zeros((na, g.H, g.W), np.int32)
    lg = np.zeros((na, g.H, g.W))
    vu = np.ascontiguousarray(vol) if vol.dtype == np.uint16 else None
    vf = np.ascontiguousarray(vol, np.float32) if vol.dtype != np.uint16 else None
    gc = g.c()
    lib().orc_line_integrals(C.byref(gc), _p(cs, _D), _p(sn, _D), _p(xc, _D), _p(yc, _D), _p(vu, _U16),
                             _p(vf, _F), 0, na, _p(L, _D), _p(ns, _I32), _p(lg, _D))
    return L, ns, lg


def forward_collapsed(g: Geometry, vol, atten, pdf, xfactor):
    vu = np.ascontiguousarray(vol) if vol.dtype == np.uint16 else None
    vf = np.ascontiguousarray(vol, np.float32) if vol.dtype != np.uint16 else None
    at = np.ascontiguousarray(atten, np.float32)
    pd = np.ascontiguousarray(pdf, np.float32)
    out = np.zeros((g.n_ang, g.H, g.W), np.float32)
    cs, sn, xc, yc = g.tables()
    gc = g.c()
    lib().orc_forward_collapsed(C.byref(gc), _p(cs, _D), _p(sn, _D), _p(xc, _D), _p(yc, _D), _p(vu, _U16),
                                _p(vf, _F), _p(at, _F), _p(pd, _F), len(pd), C.c_double(xfactor), 0, g.n_ang,
                                _p(out, _F))
    return out


def logproj(g: Geometry, vol_f32, atten, pdf, xfactor):
    vf = np.ascontiguousarray(vol_f32, np.float32)
    at = np.ascontiguousarray(atten, np.float32)
    pd = np.ascontiguousarray(pdf, np.float32)
    p = np.zeros((g.n_ang, g.H, g.W))
    cs, sn, xc, yc = g.tables()
    gc = g.c()
    lib().orc_logproj(C.byref(gc), _p(cs, _D), _p(sn, _D), _p(xc, _D), _p(yc, _D), _p(vf, _F), _p(at, _F),
                      _p(pd, _F), len(pd), C.c_double(xfactor), _p(p, _D))
    return p


def adjoint(g: Geometry, vol_f32, atten, pdf, xfactor, target):
    vf = np.ascontiguousarray(vol_f32, np.float32)
    at = np.ascontiguousarray(atten, np.float32)
    pd = np.ascontiguousarray(pdf, np.float32)
    tg = np.ascontiguousarray(target, np.float64)
    grad = np.zeros(vf.shape)
    loss = C.c_double(0)
    cs, sn, xc, yc = g.tables()
    gc = g.c()
    lib().orc_adjoint(C.byref(gc), _p(cs, _D), _p(sn, _D), _p(xc, _D), _p(yc, _D), _p(vf, _F), _p(at, _F),
                      _p(pd, _F), len(pd), C.c_double(xfactor), _p(tg, _D), _p(grad, _D), C.byref(loss))
    return grad, loss.value


def const_map(W, pitch_iso=0.032509752, alpha=0.5):
    """prepareConstMap's matrix; alpha: the window parameter (0.5 the reference's, 1 the band-limited ramp)."""
    out = np.zeros((W, W))
    if alpha == 0.5:
        lib().orc_const_map(W, C.c_double(pitch_iso), _p(out, _D))
    else:
        lib().orc_const_map_w(W, C.c_double(pitch_iso), C.c_double(alpha), _p(out, _D))
    return out


def filter_sinogram(proj, xshift, yshift, pitch_iso=0.032509752, acc64=False, alpha=0.5):
    n, H, W = proj.shape
    p = np.ascontiguousarray(proj, np.float32)
    cm = const_map(W, pitch_iso, alpha)
    xs = np.ascontiguousarray(xshift, np.float64)
    ys = np.ascontiguousarray(yshift, np.float64)
    if alpha != 0.5:  # the window option exists in the FP64 form only
        out = np.zeros((n, H, W))
        lib().orc_filter_sinogram64_w(H, W, n, _p(p, _F), _p(out, _D), _p(cm, _D), _p(xs, _D), _p(ys, _D),
                                      C.c_double(pitch_iso), C.c_double(alpha))
    elif acc64:
        out = np.zeros((n, H, W))
        lib().orc_filter_sinogram64(H, W, n, _p(p, _F), _p(out, _D), _p(cm, _D), _p(xs, _D), _p(ys, _D),
                                    C.c_double(pitch_iso))
    else:
        out = np.zeros((n, H, W), np.float32)
        lib().orc_filter_sinogram(H, W, n, _p(p, _F), _p(out, _F), _p(cm, _D), _p(xs, _D), _p(ys, _D),
                                  C.c_double(pitch_iso))
    return out


def mask(R):
    m = np.zeros((R, R), np.int8)
    lib().orc_mask(R, m.ctypes.data_as(_CH))
    return m


def reconstruct_image(R, Z, zstart, step, scale, sino, xfactor, angles, xshift, pitch_iso=0.032509752):
    n, H, W = sino.shape
    s = np.ascontiguousarray(sino, np.float32)
    ang = np.ascontiguousarray(angles, np.float64)
    cs = np.cos(ang)
    sn = np.sin(ang)
    lib().orc_trig_table_fdk(n, _p(ang, _D), _p(cs, _D), _p(sn, _D))
    xs = np.ascontiguousarray(xshift, np.float64)
    m = mask(R)
    out = np.zeros((Z, R, R))
    lib().orc_reconstruct_image(R, Z, int(zstart), C.c_double(step), C.c_double(scale), _p(out, _D),
                                m.ctypes.data_as(_CH), H, W, n, _p(s, _F), int(xfactor), _p(ang, _D),
                                _p(cs, _D), _p(sn, _D), _p(xs, _D), C.c_double(pitch_iso))
    return out


def fdk_angles(flexmap, total, tgt_class="*"):
    a = np.zeros(max(total, 1))
    x = np.zeros(max(total, 1))
    y = np.zeros(max(total, 1))
    n = lib().orc_fdk_angles(flexmap.encode(), tgt_class.encode(), total, _p(a, _D), _p(x, _D), _p(y, _D), total)
    if n < 0:
        raise RuntimeError(f"orc_fdk_angles failed: {n}")
    return a[:n], x[:n], y[:n]


# --------------------------------------------------------------------------------------------------
# The reference's own CPU code (built from its sources into oracle/_ref by oracle/Makefile).

_ref = {}


def ref_available():
    return all(os.path.exists(os.path.join(HERE, "_ref", f)) for f in ("libref_proj.so", "libref_fdk.so"))


def ref(name):
    if name not in _ref:
        path = os.path.join(HERE, "_ref", f"libref_{name}.so")
        if not os.path.exists(path):
            raise RuntimeError(f"reference build missing: {path}")
        L = C.CDLL(path)
        if name == "proj":
            L.ref_reprojection_timed.restype = C.c_double
        else:
            L.ref_filter_sinogram.restype = C.c_double
            L.ref_reconstruct_image.restype = C.c_double
        _ref[name] = L
    return _ref[name]


def ref_attenuation_coefficient(spectrum_csv, xraydb_parent, nmat=6, nthin=1, mat_type="human"):
    atten = np.zeros(1024 * nmat, np.float32)
    pdf = np.zeros(1024, np.float32)
    ne = ref("proj").ref_attenuation_coefficient(os.path.abspath(spectrum_csv).encode(),
                                                   xraydb_parent.encode(), nmat, nthin, mat_type.encode(),
                                                   _p(atten, _F), _p(pdf, _F))
    if ne <= 0:
        raise RuntimeError(f"ref attenuation_coefficient failed: {ne}")
    return atten[: nmat * ne].reshape(nmat, ne).copy(), pdf[:ne].copy()


def ref_trig_table(W, H, angles, xs, ys, det_pitch=0.05, pixnum=800.0):
    n = len(angles)
    a, x, y = [np.ascontiguousarray(v, np.float64) for v in (angles, xs, ys)]
    out = [np.zeros(n) for _ in range(4)]
    ref("proj").ref_trig_table(W, H, n, _p(a, _D), _p(x, _D), _p(y, _D), C.c_double(det_pitch),
                               C.c_double(pixnum), *[_p(o, _D) for o in out])
    return out


def _ref_proj_args(g: Geometry, vol, angles, xfactor, atten, pdf):
    cs, sn, xc, yc = g.tables()
    return dict(vol=np.ascontiguousarray(vol, np.uint16), ang=np.ascontiguousarray(angles, np.float64),
                cs=cs, sn=sn, xc=xc, yc=yc, at=np.ascontiguousarray(atten, np.float32).ravel(),
                pd=np.ascontiguousarray(pdf, np.float32))


def ref_reprojection_one(g: Geometry, vol, angles, xfactor, atten, pdf, a, ke):
    """The CPU twin for angle a, bin ke: [H][W] un-flipped, no cos^3 factor."""
    A = _ref_proj_args(g, vol, angles, xfactor, atten, pdf)
    out = np.zeros((g.H, g.W), np.float32)
    ref("proj").ref_reprojection_one(
        g.N, g.Nz, _p(A["vol"], _U16), C.c_double(g.s), C.c_double(g.t), g.W, g.H, a, _p(A["ang"], _D),
        _p(A["cs"], _D), _p(A["sn"], _D), _p(A["xc"], _D), _p(A["yc"], _D), C.c_double(xfactor),
        C.c_double(g.det_pitch), C.c_double(g.det_pitch_iso), C.c_double(g.sad), C.c_double(g.det_pixel_num),
        C.c_double(g.x_reg), C.c_double(g.y_reg), C.c_double(g.z_reg), _p(A["at"], _F), _p(A["pd"], _F),
        g.nmat, len(A["pd"]), ke, _p(out, _F))
    return out


def ref_reprojection_timed(g: Geometry, vol, angles, xfactor, atten, pdf, pairs, nthreads):
    """Times the CPU twin over (angle, bin) pairs on nthreads host threads.  Returns (seconds, out)."""
    A = _ref_proj_args(g, vol, angles, xfactor, atten, pdf)
    pa = np.ascontiguousarray([p[0] for p in pairs], np.int32)
    pk = np.ascontiguousarray([p[1] for p in pairs], np.int32)
    out = np.zeros((len(pairs), g.H, g.W), np.float32)
    sec = ref("proj").ref_reprojection_timed(
        g.N, g.Nz, _p(A["vol"], _U16), C.c_double(g.s), C.c_double(g.t), g.W, g.H, _p(A["ang"], _D),
        _p(A["cs"], _D), _p(A["sn"], _D), _p(A["xc"], _D), _p(A["yc"], _D), C.c_double(xfactor),
        C.c_double(g.det_pitch), C.c_double(g.det_pitch_iso), C.c_double(g.sad), C.c_double(g.det_pixel_num),
        C.c_double(g.x_reg), C.c_double(g.y_reg), C.c_double(g.z_reg), _p(A["at"], _F), _p(A["pd"], _F),
        g.nmat, len(A["pd"]), len(pairs), _p(pa, _I32), _p(pk, _I32), int(nthreads), _p(out, _F))
    return sec, out


def ref_const_map(W, pitch=0.032509752):
    out = np.zeros((W, W))
    ref("fdk").ref_const_map(W, C.c_double(pitch), _p(out, _D))
    return out


def ref_filter_sinogram(proj, xshift, yshift, pitch=0.032509752):
    n, H, W = proj.shape
    p = np.ascontiguousarray(proj, np.float32)
    out = np.zeros_like(p)
    xs = np.ascontiguousarray(xshift, np.float64)
    ys = np.ascontiguousarray(yshift, np.float64)
    sec = ref("fdk").ref_filter_sinogram(H, W, n, _p(p, _F), _p(out, _F), _p(xs, _D), _p(ys, _D), C.c_double(pitch))
    return out, sec


def ref_reconstruct_image(R, Z, zstart, step, scale, sino, xfactor, angles, xshift, pitch=0.032509752):
    n, H, W = sino.shape
    s = np.ascontiguousarray(sino, np.float32)
    out = np.zeros((Z, R, R))
    a = np.ascontiguousarray(angles, np.float64)
    xs = np.ascontiguousarray(xshift, np.float64)
    sec = ref("fdk").ref_reconstruct_image(R, Z, int(zstart), C.c_double(step), C.c_double(scale), _p(out, _D),
                                           H, W, n, _p(s, _F), int(xfactor), _p(a, _D), _p(xs, _D),
                                           C.c_double(pitch))
    return out, sec


def ref_mask(R):
    m = np.zeros((R, R), np.int8)
    ref("fdk").ref_mask(R, m.ctypes.data_as(_CH))
    return m


def ref_fdk_load(txt, workdir, cap=4096):
    ints = np.zeros(8, np.int32)
    dbls = np.zeros(3)
    strs = C.create_string_buffer(1280)
    a, x, y = np.zeros(cap), np.zeros(cap), np.zeros(cap)
    n = ref("fdk").ref_fdk_load(txt.encode(), workdir.encode(), _p(ints, _I32), _p(dbls, _D), strs,
                                _p(a, _D), _p(x, _D), _p(y, _D), cap)
    if n < 0:
        raise RuntimeError(f"ref_fdk_load failed {n}")
    raw = strs.raw
    s = [raw[i * 256:(i + 1) * 256].split(b"\0")[0].decode() for i in range(5)]
    return dict(ints=ints, dbls=dbls, strs=s[:4], size_class=s[4], angles=a[:n], xshift=x[:n], yshift=y[:n])


def ref_proj_load(txt, workdir, cap=4096, vol_cap=0):
    ints = np.zeros(9, np.int32)
    dbls = np.zeros(7)
    strs = C.create_string_buffer(1024)
    a, x, y = np.zeros(cap), np.zeros(cap), np.zeros(cap)
    vol = np.zeros(max(vol_cap, 1), np.uint16)
    n = ref("proj").ref_proj_load(txt.encode(), workdir.encode(), _p(ints, _I32), _p(dbls, _D), strs,
                                  _p(a, _D), _p(x, _D), _p(y, _D), cap, _p(vol, _U16) if vol_cap else None,
                                  C.c_long(vol_cap))
    if n < 0:
        raise RuntimeError(f"ref_proj_load failed {n}")
    raw = strs.raw
    s = [raw[i * 256:(i + 1) * 256].split(b"\0")[0].decode() for i in range(4)]
    return dict(ints=ints, dbls=dbls, strs=s, angles=a[:n], xshift=x[:n], yshift=y[:n], vol=vol)


# --------------------------------------------------------------------------------------------------
# Organ-label phantom path (ICRPphantoms_2mmslice/To_material_distribution_for512_random.cpp).

def rand_sequence(seed, n):
    out = np.zeros(n, np.int32)
    lib().orc_rand_sequence(C.c_uint(seed), n, _p(out, _I32))
    return out


def material_weights(media_csv, randomize=True, seed=1, cap=256):
    """Per-medium weight % [n][6] (H C N O P Ca) after the reference's draws."""
    w = np.zeros((cap, 6), np.float32)
    n = lib().orc_material_weights(os.fspath(media_csv).encode(), int(bool(randomize)), C.c_uint(seed), _p(w, _F), cap)
    if n < 0:
        raise RuntimeError(f"orc_material_weights failed ({n})")
    return w[:n].copy()


def organ_table(organs_csv, cap=1024):
    organ = np.zeros(cap, np.int32)
    medium = np.zeros(cap, np.int32)
    dens = np.zeros(cap, np.float32)
    n = lib().orc_organ_table(os.fspath(organs_csv).encode(), _p(organ, _I32), _p(medium, _I32), _p(dens, _F), cap)
    if n < 0:
        raise RuntimeError(f"orc_organ_table failed ({n})")
    return organ[:n].copy(), medium[:n].copy(), dens[:n].copy()


def labels_to_maps(labels, weights, organs):
    """labels uint8 [nz][nrow][ncol] -> int16 [nz][6][nrow][ncol] (the reference's voxel loop)."""
    lab = np.ascontiguousarray(labels, np.uint8)
    nz, nrow, ncol = lab.shape
    w = np.ascontiguousarray(weights, np.float32)
    organ, medium, dens = (np.ascontiguousarray(a) for a in organs)
    out = np.empty((nz, 6, nrow, ncol), np.int16)
    rc = lib().orc_labels_to_maps(lab.ctypes.data_as(C.POINTER(C.c_uint8)), ncol, nrow, nz, _p(w, _F), len(w),
                                  _p(organ, _I32), _p(medium, _I32), _p(dens, _F), len(organ),
                                  out.ctypes.data_as(C.POINTER(C.c_int16)))
    if rc:
        raise RuntimeError(f"orc_labels_to_maps failed ({rc})")
    return out


def material_lut(media_csv, organs_csv, randomize=True, seed=1):
    """[256][6] int16: the voxel loop applied to the labels 0..255."""
    w = material_weights(media_csv, randomize, seed)
    lab = np.arange(256, dtype=np.uint8).reshape(1, 1, 256)
    return labels_to_maps(lab, w, organ_table(organs_csv))[0, :, 0, :].T.copy()


REF_TO_MATERIAL = os.path.join(HERE, "_ref", "to_material")
REF_TO_MATERIAL_SHAPE = (100, 271, 543)  # slices, rows, columns: fixed in the reference program


def ref_to_material_available():
    return os.path.exists(REF_TO_MATERIAL)


def ref_to_material_run(labels, media_csv, organs_csv, workdir):
    """Runs the reference converter (built from its source into oracle/_ref) on a label volume of its fixed
    shape in `workdir`.  Returns (test.raw as int16 [6][100][271][543], the per-slice maps as int16
    [100][6][271][543], stdout text)."""
    import shutil
    import subprocess
    lab = np.ascontiguousarray(labels, np.uint8)
    if lab.shape != REF_TO_MATERIAL_SHAPE:
        raise ValueError(f"the reference program reads exactly {REF_TO_MATERIAL_SHAPE} labels")
    os.makedirs(os.path.join(workdir, "AM_MD"), exist_ok=True)
    os.makedirs(os.path.join(workdir, "AM_ED"), exist_ok=True)
    lab.tofile(os.path.join(workdir, "AM_skin_543x271_700_800.raw"))
    shutil.copyfile(media_csv, os.path.join(workdir, "AM_media.csv"))
    shutil.copyfile(organs_csv, os.path.join(workdir, "AM_organs.csv"))
    r = subprocess.run([REF_TO_MATERIAL], cwd=workdir, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"reference converter failed: {r.stderr[-500:]}")
    nz, nrow, ncol = REF_TO_MATERIAL_SHAPE
    whole = np.fromfile(os.path.join(workdir, "test.raw"), np.int16).reshape(6, nz, nrow, ncol)
    md = np.empty((nz, 6, nrow, ncol), np.int16)
    for iz in range(nz):
        md[iz] = np.fromfile(os.path.join(workdir, "AM_MD", f"AM_MD_{ncol}_{nrow}_{iz:03d}.raw"),
                             np.int16).reshape(6, nrow, ncol)
    return whole, md, r.stdout


# --------------------------------------------------------------------------------------------------
# Iterative-reconstruction pieces (P/lengthproj.cpp, P/IR_ImageCBCT.cpp).

def lengthproj(g: Geometry, proj_u16, ite=0, a0=0, a1=None):
    """npf [Nz][N][N] FP32: chord-weighted voxel-driven backprojection of (65535 - proj)."""
    p = np.ascontiguousarray(proj_u16, np.uint16)
    a1 = g.n_ang if a1 is None else a1
    cs, sn, xc, yc = g.tables()
    out = np.empty((g.Nz, g.N, g.N), np.float32)
    lib().orc_lengthproj(C.byref(g.c()), int(ite), _p(cs, _D), _p(sn, _D), _p(xc, _D), _p(yc, _D), _p(p, _U16),
                         int(a0), int(a1), _p(out, _F))
    return out


def ir_update(before_i16, npf, npf_re, scale, wbefore, weight_tv, weight_prior=0.0):
    b = np.ascontiguousarray(before_i16, np.int16)
    Nz, N, _ = b.shape
    out = np.empty(b.shape, np.float32)
    lib().orc_ir_update(N, Nz, C.c_double(scale), b.ctypes.data_as(C.POINTER(C.c_int16)),
                        _p(np.ascontiguousarray(npf, np.float32), _F), _p(np.ascontiguousarray(npf_re, np.float32), _F),
                        C.c_double(wbefore), C.c_double(weight_tv), C.c_double(weight_prior), _p(out, _F))
    return out


def ref_lengthproj(g: Geometry, proj_u16, angles, ite=0, a0=0, a1=None):
    """The reference's lengthproj (built from its sources into oracle/_ref)."""
    p = np.ascontiguousarray(proj_u16, np.uint16)
    a1 = g.n_ang if a1 is None else a1
    cs, sn, xc, yc = g.tables()
    ang = np.ascontiguousarray(angles, np.float64)
    out = np.empty((g.Nz, g.N, g.N), np.float32)
    d = C.c_double
    ref("proj").ref_lengthproj(int(ite), g.N, g.Nz, _p(p, _U16), _p(out, _F), d(g.s), d(g.t), g.W, g.H, int(a0),
                               int(a1), _p(ang, _D), _p(cs, _D), _p(sn, _D), _p(xc, _D), _p(yc, _D), d(g.det_pitch),
                               d(g.det_pitch_iso), d(g.sad), d(g.det_pixel_num), d(g.x_reg), d(g.y_reg), d(g.z_reg))
    return out


def ref_ir_update(before_i16, npf, npf_re, scale, wbefore, weight_tv, weight_prior=0.0):
    b = np.ascontiguousarray(before_i16, np.int16)
    Nz, N, _ = b.shape
    out = np.empty(b.shape, np.float32)
    d = C.c_double
    ref("proj").ref_ir_update(N, Nz, d(scale), b.ctypes.data_as(C.POINTER(C.c_int16)), _p(out, _F),
                              _p(np.ascontiguousarray(npf, np.float32), _F),
                              _p(np.ascontiguousarray(npf_re, np.float32), _F), d(wbefore), d(weight_tv), d(weight_prior))
    return out


# --------------------------------------------------------------------------------------------------
# FDK angle mask (F/misc.cpp:87-180) and the panel-offset class (F/projectionDataMVkV.cpp:277-281).

def size_class(xshift):
    out = C.create_string_buffer(4)
    x = np.ascontiguousarray(xshift, np.float64)
    lib().orc_size_class(len(x), _p(x, _D), out)
    return out.value.decode()


def angle_mask(angles_rad, R, scale, cls):
    a = np.ascontiguousarray(angles_rad, np.float64)
    m = np.empty((len(a), R, R), np.uint8)
    lib().orc_angle_mask(len(a), _p(a, _D), R, C.c_double(scale), cls.encode(), m.ctypes.data_as(_CH))
    return m


def offset_exponents(W, xshift, taper=0.0):
    """[n][W] exponents 2 w of the offset-detector weighting (orc_offset_exponent); -1 where no overlap."""
    xs = np.atleast_1d(np.asarray(xshift, np.float64))
    f = lib().orc_offset_exponent
    f.restype = C.c_double
    return np.array([[f(int(W), C.c_double(x), C.c_double(taper), c) for c in range(W)] for x in xs])


def half_fan_width(W, xshift):
    xs = np.ascontiguousarray(xshift, np.float64)
    return int(lib().orc_half_fan_width(int(W), len(xs), _p(xs, _D)))


def offset_weight(p, xshift, yshift, taper=0.0):
    """(padded panel [n][H][W + P], xshift', yshift') of the half-fan weighting (orc_offset_weight)."""
    p = np.ascontiguousarray(p, np.float32)
    n, H, W = p.shape
    xs = np.ascontiguousarray(xshift, np.float64)
    ys = np.ascontiguousarray(yshift, np.float64)
    Wo = half_fan_width(W, xs)
    if Wo < 0:
        raise ValueError("offset weighting: shifts of both signs, or one leaves no overlap")
    out = np.empty((n, H, Wo), np.float32)
    xo, yo = np.empty(n), np.empty(n)
    rc = lib().orc_offset_weight(n, H, W, _p(p, _F), _p(xs, _D), _p(ys, _D), C.c_double(taper), Wo, _p(out, _F),
                                 _p(xo, _D), _p(yo, _D))
    if rc != 0:
        raise ValueError("offset weighting failed")
    return out, xo, yo


def philox4x32_10(ctr, key):
    """Philox4x32-10 of one 4-word counter under a 2-word key (orc_philox4x32_10)."""
    c = (C.c_uint32 * 4)(*[int(x) & 0xFFFFFFFF for x in ctr])
    k = (C.c_uint32 * 2)(*[int(x) & 0xFFFFFFFF for x in key])
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return [int(x) for x in o]


def log_factorial(k):
    """log k! as the detector model's sampler forms it (orc_log_factorial)."""
    f = lib().orc_log_factorial
    f.restype = C.c_double
    f.argtypes = [C.c_double]
    return float(f(float(k)))


def poisson(lam, index, seed):
    """The detector model's Poisson draw for one pixel (orc_poisson)."""
    f = lib().orc_poisson
    f.restype = C.c_longlong
    f.argtypes = [C.c_double, C.c_uint64, C.c_uint64]
    return int(f(float(lam), int(index), int(seed)))


def detector_counts(primary, scatter, I0, seed, raw=False):
    """Poisson counts of I0 (primary + scatter) per pixel, raw or as max(k, 1) / I0 (orc_detector_counts)."""
    p = np.ascontiguousarray(primary, np.float32)
    s = None if scatter is None else np.ascontiguousarray(scatter, np.float32)
    out = np.empty_like(p)
    f = lib().orc_detector_counts
    f.argtypes = [C.c_size_t, _F, _F, C.c_double, C.c_uint64, C.c_int, _F]
    f(p.size, _p(p, _F), _p(s, _F), float(I0), int(seed), int(bool(raw)), _p(out, _F))
    return out


def ref_angle_mask(angles_rad, R, scale, cls):
    a = np.ascontiguousarray(angles_rad, np.float64)
    m = np.empty((len(a), R, R), np.uint8)
    ref("fdk").ref_prepare_angles(len(a), _p(a, _D), R, C.c_double(scale), cls.encode(), m.ctypes.data_as(_CH))
    return m
