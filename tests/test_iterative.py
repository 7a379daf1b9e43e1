"""The iterative-reconstruction pieces (P/lengthproj.cpp, P/IR_ImageCBCT.cpp): the FP64 oracle against the
reference's own functions compiled from its sources (bit-exact), and against the committed fixture of their
outputs."""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "iterative_small.npz")
need_ref = pytest.mark.skipif(not O.ref_available(), reason="reference CPU code not built (oracle/_ref)")


def geom(N, Nz, s, t, H, W, angles_deg, xs=0.013, ys=-0.02, x_reg=0.0, y_reg=0.0, z_reg=0.0):
    ang = np.deg2rad(np.asarray(angles_deg, float))
    cs, sn, xc, yc = O.trig_table_proj(W, H, ang, np.full(len(ang), xs), np.full(len(ang), ys))
    return O.Geometry(N=N, s=s, Nz=Nz, t=t, H=H, W=W, cos_t=cs, sin_t=sn, xc=xc, yc=yc, x_reg=x_reg, y_reg=y_reg,
                      z_reg=z_reg), ang


LP_CASES = [
    (24, 6, 0.5, 0.4, 20, 28, [-170.0001, -90.00001, 0.0, 33.3, 90.0, 179.99999], 0.0, 0.0, 0.0),
    (17, 5, 0.3, 0.7, 31, 25, [-45.0, 10.0, 120.5, 200.0], 0.3, 0.1, -0.2),   # odd sizes, W < H, registration
    (32, 4, 1.1, 0.2, 16, 16, list(np.arange(-180, 180, 20.0) + 0.5), 0.0, 0.0, 0.9),  # wide voxels, taps off the edge
]


@need_ref
@pytest.mark.parametrize("case", LP_CASES)
@pytest.mark.parametrize("ite", [0, 1])
def test_lengthproj_oracle_matches_reference(case, ite):
    N, Nz, s, t, H, W, angs, xr, yr, zr = case
    g, ang = geom(N, Nz, s, t, H, W, angs, x_reg=xr, y_reg=yr, z_reg=zr)
    rng = np.random.default_rng(N + Nz + ite)
    proj = rng.integers(0, 65536, (len(angs), H, W)).astype(np.uint16)
    a = O.lengthproj(g, proj, ite)
    b = O.ref_lengthproj(g, proj, ang, ite)
    assert np.array_equal(a, b)
    assert (a != 0).mean() > 0.3
    # a sub-range of angles
    assert np.array_equal(O.lengthproj(g, proj, ite, 1, 3), O.ref_lengthproj(g, proj, ang, ite, 1, 3))


@need_ref
def test_lengthproj_angle_zero_contributes_nothing():
    """sin = 0 makes the detector-line slope infinite; on the reference's build the sample index becomes INT_MIN
    and the angle is dropped for every voxel."""
    g, ang = geom(16, 3, 0.5, 0.5, 12, 12, [0.0])
    proj = np.full((1, 12, 12), 1000, np.uint16)
    assert not O.ref_lengthproj(g, proj, ang).any()
    assert not O.lengthproj(g, proj).any()


def test_lengthproj_closed_form_uniform_projection():
    """With a constant projection every accepted sample is the constant, so npf = (65535 - c) x (sum of chords
    of the accepted angles); for a voxel at the centre of a central slice, the chord of an angle whose source
    lies on an axis is the voxel width (the line runs along that axis)."""
    g, ang = geom(8, 2, 0.5, 0.5, 64, 64, [90.0, 180.0, -90.0])
    proj = np.full((3, 64, 64), 5535, np.uint16)
    npf = O.lengthproj(g, proj)
    # voxel (x=4, y=3) has centre (0.25, 0.25): source at (-100 sin, 100 cos); all three angles sample inside
    # the 64 x 64 detector; the chord is the x-extent (or y-extent) times sec of the small tilt
    c = npf[0, 3, 4] / 60000.0
    assert 3 * 0.5 <= c < 3 * 0.5 * 1.001


@need_ref
def test_ir_update_oracle_matches_reference():
    rng = np.random.default_rng(5)
    for N, Nz, s in ((24, 6, 0.5), (13, 4, 1.3), (30, 3, 0.1)):
        before = rng.integers(-50, 3000, (Nz, N, N)).astype(np.int16)
        before[0, 1, 1] = 32000  # ratio large enough to clamp at 5000; its neighbour's int TV sum overflows to NaN
        npf = rng.uniform(1e3, 1e6, (Nz, N, N)).astype(np.float32)
        npr = rng.uniform(1e3, 1e6, (Nz, N, N)).astype(np.float32)
        npr[0, 1, 1] = 1e7
        for w, tv, pr in ((0.3, 0.05, 0.0), (0.0, 0.0, 0.0), (0.9, 2.0, 1.0)):
            a = O.ir_update(before, npf, npr, s, w, tv, pr)
            b = O.ref_ir_update(before, npf, npr, s, w, tv, pr)
            assert np.array_equal(a, b, equal_nan=True), (N, Nz, s, w, tv)


def test_ir_update_closed_form():
    """Without TV (weight 0) the update is x (w + (1 - w) r / (n + 1e-6)) clamped; flat images have zero TV."""
    N, Nz = 10, 4
    before = np.full((Nz, N, N), 1200, np.int16)
    npf = np.full((Nz, N, N), 2.0, np.float32)
    npr = np.full((Nz, N, N), 3.0, np.float32)
    out = O.ir_update(before, npf, npr, 0.5, 0.25, 7.0)
    expect = np.float32(1200 * (0.25 + 0.75 * 3.0 / (2.0 + 0.000001)))
    assert np.all(out == expect)


def test_iterative_fixture():
    """The reference's outputs committed as a fixture (tests/golden/gen_golden.py)."""
    d = np.load(GOLD)
    g = O.Geometry(N=int(d["N"]), s=float(d["s"]), Nz=int(d["Nz"]), t=float(d["t"]), H=int(d["H"]), W=int(d["W"]),
                   cos_t=d["cos"], sin_t=d["sin"], xc=d["xc"], yc=d["yc"], x_reg=float(d["x_reg"]),
                   z_reg=float(d["z_reg"]))
    assert np.array_equal(O.lengthproj(g, d["proj"], 0), d["npf"])
    assert np.array_equal(O.lengthproj(g, d["proj"], 1), d["npf_ite1"])
    assert np.array_equal(O.ir_update(d["before"], d["npf"], d["npf_re"], float(d["s"]), 0.3, 0.05), d["ir"])


# ------------------------------------------------------------- FDK angle mask and panel-offset class

@need_ref
@pytest.mark.parametrize("R", [16, 17, 64])
@pytest.mark.parametrize("cls", ["S", "M", "L"])
def test_angle_mask_oracle_matches_reference(R, cls):
    rng = np.random.default_rng(R)
    ang = np.concatenate([np.sort(rng.uniform(-np.pi, np.pi, 30)), [-np.pi, np.pi, 0.0, np.pi / 2]])
    assert np.array_equal(O.angle_mask(ang, R, 0.1, cls), O.ref_angle_mask(ang, R, 0.1, cls))


def test_angle_mask_closed_form():
    """A voxel near the +y axis (th1 = -atan(x/y), -3 deg) keeps the 200-degree window from th1 + 170 deg up
    through +-180 deg to th1 + 10 deg and drops the rest; the mask of an S panel is all ones."""
    R = 20
    ang = np.deg2rad(np.array([-169.0, -90.0, 0.0, 6.0, 8.0, 90.0, 166.0, 168.0, 175.0]))
    m = O.angle_mask(ang, R, 0.1, "M")
    j, i = 0, R // 2  # x = 0.05, y = 0.95 (scale units)
    th1 = -np.arctan(0.5 / 9.5)
    ts, te = th1 + np.pi - np.pi / 18, th1 + np.pi / 18
    expect = [bool(a > ts or a < te) for a in ang]
    assert expect == [True, True, True, True, False, False, False, True, True]
    assert list(m[:, j, i].astype(bool)) == expect
    assert O.angle_mask(ang, R, 0.1, "S").all()


def test_size_class_thresholds():
    assert O.size_class([5.9] * 4) == "S" and O.size_class([6.0, 6.0]) == "M"
    assert O.size_class([15.0]) == "M" and O.size_class([15.1, 15.1]) == "L"
    from virtualcbct_amd import fdk
    for xs in ([5.9] * 4, [6.0, 6.0], [15.0], [15.1, 15.1], [0.0, 30.0]):
        assert fdk.size_class(xs) == O.size_class(xs)


@need_ref
def test_size_class_matches_reference_loader(tmp_path, ref_dir):
    """The FDK loader's class for the shipped 36-angle txt (its flexmap shifts)."""
    F = f"{ref_dir}/Release_recon_from_forward_pro_float_CM"
    d = O.ref_fdk_load(f"{F}/txtfile/FDK_36pro_water20cm.txt", F)
    assert d["size_class"] == O.size_class(d["xshift"])
