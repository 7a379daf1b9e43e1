"""The GPU kernels' traversal (virtualcbct_amd/csrc/traverse.h), compiled for the host and run with the same
64-row wave grouping and wave-wide step culling as the kernel, against the oracle's literal restatement of
reprojectionKernel.  Catches logic errors in the lockstep walk before any GPU run."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import oracle as O
from virtualcbct_amd import phantoms

HERE = os.path.join(os.path.dirname(__file__), "host_traverse")
LIB = os.path.join(HERE, "libhosttrav.so")
LIBS = {"plain": os.path.join(HERE, "libhosttrav.so"), "fma": os.path.join(HERE, "libhosttrav_fma.so")}
D = C.POINTER(C.c_double)


def host_line_integrals(g, vol, lib=None, table=True):
    lib = lib or LIB
    if not os.path.exists(lib):
        pytest.skip(f"host traversal harness not built: {lib}")
    L_ = C.CDLL(lib)
    L_.ht_set_table(1 if table else 0)
    cs, sn, xc, yc = g.tables()
    tab = np.ascontiguousarray(np.concatenate([cs, sn, xc, yc]))
    L = np.zeros((g.n_ang, g.H, g.W, g.nmat))
    ns = np.zeros((g.n_ang, g.H, g.W), np.int32)
    v = np.ascontiguousarray(vol, np.uint16)
    L_.ht_line_integrals(g.N, C.c_double(g.s), g.Nz, C.c_double(g.t), g.H, g.W, g.n_ang, C.c_double(g.x_reg),
                         C.c_double(g.y_reg), C.c_double(g.z_reg), C.c_double(g.det_pitch), C.c_double(g.det_pitch_iso),
                         C.c_double(g.sad), C.c_double(g.det_pixel_num), g.nmat, tab.ctypes.data_as(D),
                         v.ctypes.data_as(C.POINTER(C.c_uint16)), L.ctypes.data_as(D),
                         ns.ctypes.data_as(C.POINTER(C.c_int32)))
    return L, ns


def geom(N, Nz, s, t, H, W, angles_deg, xs=0.0, z_reg=0.0):
    ang = np.deg2rad(np.asarray(angles_deg, float))
    cs, sn, xc, yc = O.trig_table_proj(W, H, ang, np.full(len(ang), xs), np.zeros(len(ang)))
    return O.Geometry(N=N, s=s, Nz=Nz, t=t, H=H, W=W, cos_t=cs, sin_t=sn, xc=xc, yc=yc, z_reg=z_reg)


ANGLES = [-170.0001, -90.00001, -40.00001, 9.99999, 45.0, 59.99999, 133.3, 179.99999]


@pytest.mark.parametrize("table", [True, False], ids=["step-table", "per-lane"])
@pytest.mark.parametrize("build", ["plain", "fma"])
@pytest.mark.parametrize("case", [
    (32, 8, 0.5, 0.5, 24, 24, 0.3, 0.0123),
    (32, 8, 0.5, 0.5, 24, 24, 0.0, 0.0),       # centre column on the rotation axis: exact y = 0 ties
    (64, 16, 0.25, 0.3, 80, 70, -0.7, 0.3),    # z_reg shift, non-square detector
    (48, 100, 0.3, 0.05, 130, 20, 0.0, -1.6923),  # many thin slices: many z crossings per ray (the panel
    #                                              shift puts the central ray, at column H/2 + shift, on the panel)
    (41, 9, 0.37, 0.21, 33, 29, 0.11, -0.2),   # odd sizes, box half-width not on the plane grid
    (16, 4, 0.8, 1.2, 14, 14, 0.0, 0.01),      # planes offset by s/2 from the voxel grid: every midpoint a tie
])
def test_host_traversal_matches_oracle(case, build, table):
    N, Nz, s, t, H, W, zr, xs = case
    rng = np.random.default_rng(N * 1000 + Nz)
    g = geom(N, Nz, s, t, H, W, ANGLES, xs, zr)
    vol = rng.integers(0, 2000, (Nz, 6, N, N)).astype(np.uint16)
    Lo, _, _ = O.line_integrals(g, vol)
    Lh, _ = host_line_integrals(g, vol, LIBS[build], table)
    den = np.maximum(np.abs(Lo).sum(-1), 1e-9)
    err = np.abs(Lo - Lh).sum(-1) / den
    assert err.max() < 1e-8, np.unravel_index(err.argmax(), err.shape)


@pytest.mark.parametrize("table", [True, False], ids=["step-table", "per-lane"])
@pytest.mark.parametrize("build", ["plain", "fma"])
def test_host_traversal_random_geometries(build, table):
    # seeded sweep of small geometries: pitches that do and do not divide 1 cm, offsets, z shifts, odd sizes
    rng = np.random.default_rng(777)
    for it in range(60):
        N = int(rng.integers(8, 40))
        s = float(rng.choice([0.05, 0.1, 0.2, 0.25, 0.3, 0.4, 0.5, 0.8, 1.0, 0.37]))
        Nz = int(rng.integers(2, 24))
        t = float(rng.choice([0.1, 0.2, 0.5, 1.0, 0.33]))
        H, W = int(rng.integers(8, 40)), int(rng.integers(8, 40))
        ang = list(rng.uniform(-180, 180, 5)) + [-90.00001, 45.0, 179.99999]
        xs = float(rng.choice([0.0, 0.013, -0.2, 1.15]))
        zr = float(rng.choice([0.0, 0.3, -0.45]))
        g = geom(N, Nz, s, t, H, W, ang, xs, zr)
        vol = rng.integers(0, 2000, (Nz, 6, N, N)).astype(np.uint16)
        Lo, _, _ = O.line_integrals(g, vol)
        Lh, _ = host_line_integrals(g, vol, LIBS[build], table)
        den = np.maximum(np.abs(Lo).sum(-1), 1e-9)
        err = np.abs(Lo - Lh).sum(-1) / den
        assert err.max() < 1e-8, (it, N, s, Nz, t, H, W, xs, zr, np.unravel_index(err.argmax(), err.shape))


def test_host_traversal_c1_geometry_one_angle():
    # full-size water cylinder of the 512^2 x 64 workload, one angle, all 160 000 rays
    w = phantoms.WORKLOADS["c1"]
    vol = phantoms.water_cylinder(w["N"], w["Nz"], w["s"], w["radius"])
    ang, xs, ys = phantoms.flexmap_angles(36, w["xshift_proj"])
    cs, sn, xc, yc = O.trig_table_proj(w["W"], w["H"], ang, xs, ys)
    sel = [7]
    g = O.Geometry(N=w["N"], s=w["s"], Nz=w["Nz"], t=w["t"], H=w["H"], W=w["W"], cos_t=cs[sel], sin_t=sn[sel],
                   xc=xc[sel], yc=yc[sel])
    Lo, _, _ = O.line_integrals(g, vol)
    Lh, _ = host_line_integrals(g, vol)
    den = np.maximum(np.abs(Lo).sum(-1), 1e-9)
    assert (np.abs(Lo - Lh).sum(-1) / den).max() < 1e-8


@pytest.mark.parametrize("build", ["plain", "fma"])
def test_host_traversal_c2_offset_geometry(build):
    # the 115 mm half-fan offset workload at two angles (one with lumped-entry ties at the cylinder edge)
    w = phantoms.WORKLOADS["c2"]
    vol = phantoms.water_cylinder(w["N"], w["Nz"], w["s"], w["radius"])
    ang, xs, ys = phantoms.flexmap_angles(360, w["xshift_proj"])
    cs, sn, xc, yc = O.trig_table_proj(w["W"], w["H"], ang, xs, ys)
    sel = [45, 200]
    g = O.Geometry(N=w["N"], s=w["s"], Nz=w["Nz"], t=w["t"], H=w["H"], W=w["W"], cos_t=cs[sel], sin_t=sn[sel],
                   xc=xc[sel], yc=yc[sel])
    Lo, _, _ = O.line_integrals(g, vol)
    Lh, _ = host_line_integrals(g, vol, LIBS[build])
    den = np.maximum(np.abs(Lo).sum(-1), 1e-9)
    assert (np.abs(Lo - Lh).sum(-1) / den).max() < 1e-8


@pytest.mark.parametrize("build", ["plain", "fma"])
@pytest.mark.parametrize("b", [100.0, 0.032509752 * 800 / 400, 0.032509752 * 800 / 630, 0.032509752 * 800 / 64,
                               0.032509752 * 800 / 33, 0.1, 3.0])
def test_division_by_known_divisor_is_exact(build, b):
    """div_known (the backprojector's divisions by the source distance and the pixel pitch): the product with
    the correctly rounded reciprocal plus one FMA correction equals the IEEE quotient on 1e6 pseudo-random
    dividends spanning 2^-30 .. 2^30 (Markstein's theorem; a dropped correction fails here)."""
    lib = LIBS[build]
    if not os.path.exists(lib):
        pytest.skip(f"host traversal harness not built: {lib}")
    f = C.CDLL(lib).ht_div_known_mismatches
    f.restype = C.c_long
    f.argtypes = [C.c_double, C.c_long, C.c_ulonglong, C.c_int, C.c_int]
    assert f(b, 1000000, 11, 30, 30) == 0


@pytest.mark.parametrize("build", ["plain", "fma"])
def test_division_with_a_shared_reciprocal_is_exact(build):
    """The same identity with a divisor drawn afresh for each quotient and its reciprocal formed by one IEEE
    division (lengthproj's shared divisors): 4e6 pseudo-random pairs over 2^-40 .. 2^40 and over 2^-4 .. 2^4."""
    lib = LIBS[build]
    if not os.path.exists(lib):
        pytest.skip(f"host traversal harness not built: {lib}")
    f = C.CDLL(lib).ht_div_known_random_mismatches
    f.restype = C.c_long
    f.argtypes = [C.c_long, C.c_ulonglong, C.c_int, C.c_int]
    assert f(2000000, 3, 40, 40) == 0
    assert f(2000000, 9, 4, 4) == 0
