"""GPU parity of the FDK filter (FFT convolution) and backprojection against the reference's own outputs
(golden fixtures) and the FP64 oracle.  Reconstruction tolerance: 1e-3 of the water level on the pre-cast
x1000 values (water ~ 200 here, so 0.2)."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from virtualcbct_amd import phantoms

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def torch():
    import torch as T
    if not T.cuda.is_available():
        pytest.skip("no GPU")
    return T


def test_filter_matches_reference_fixture(torch):
    from virtualcbct_amd import fdk
    d = np.load(os.path.join(GOLD, "fdk_small.npz"))
    for pk, xk, yk, fk in (("proj", "xshift", "yshift", "filtered"), ("proj2", "xshift12", "yshift12", "filtered2")):
        n, H, W = d[pk].shape
        out = fdk.filterSinogram(H, W, n, d[pk], d[xk], d[yk])
        ref = d[fk]
        err = np.abs(out - ref).max() / np.abs(ref).max()
        print(f"filter {pk}: max err / max|ref| = {err:.3e}")
        assert err < 2e-5


def test_filter_w400_matches_fp64_oracle(torch):
    from virtualcbct_amd import fdk
    rng = np.random.default_rng(3)
    proj = rng.uniform(0.02, 1.0, (3, 16, 400)).astype(np.float32)
    xs = np.array([0.0, 11.5, -3.0])
    ys = np.array([0.0, 0.3, 1.0])
    out = fdk.filterSinogram(16, 400, 3, proj, xs, ys)
    ref = O.filter_sinogram(proj, xs, ys, acc64=True)
    err = np.abs(out - ref).max() / np.abs(ref).max()
    print(f"filter W=400 vs FP64: {err:.3e}")
    assert err < 2e-5


@pytest.mark.parametrize("W,nrows", [(2, 3), (33, 7), (300, 6), (630, 6), (700, 9), (1100, 5), (2048, 3)])
def test_filter_other_widths_match_fp64_oracle(torch, W, nrows):
    """FFT lengths 80 = 5 x 16 (radix-4 stages and a closing radix-5 stage), 640 = 5 x 128 (a radix-2 stage
    before the radix-5 one), 1280 = 5 x 256 (the half-fan panel's 630 columns), 2048 (a power of two: 5 x 256
    is too short for 700 columns) and 2560 = 5 x 512 (1100 columns: shorter than 4096), the limits (2 columns:
    64 points; 2048 columns: 4096 points, the 64 KB of LDS a block may hold and every pixel register of a
    thread), odd row counts (the last row pair has one row) and block tails."""
    from virtualcbct_amd import fdk
    rng = np.random.default_rng(W)
    proj = rng.uniform(0.02, 1.0, (2, nrows, W)).astype(np.float32)
    xs, ys = np.array([0.4, -7.0]), np.array([0.1, 2.0])
    out = fdk.filterSinogram(nrows, W, 2, proj, xs, ys)
    ref = O.filter_sinogram(proj, xs, ys, acc64=True)
    err = np.abs(out - ref).max() / np.abs(ref).max()
    print(f"filter W={W} vs FP64: {err:.3e}")
    assert err < 2e-5


def test_backprojection_matches_reference_fixture(torch):
    from virtualcbct_amd import fdk
    d = np.load(os.path.join(GOLD, "fdk_small.npz"))
    rec = np.zeros((4, 16, 16))
    fdk.reconstructImage(16, 4, -1, 0.5, 0.3, rec, 8, 16, 8, d["sino"], 1000, d["angles"], d["xshift8"])
    err = np.abs(rec - d["recon"]).max() / np.abs(d["recon"]).max()
    assert err < 1e-5
    rec2 = np.zeros((6, 24, 24))
    fdk.reconstructImage(24, 6, -1, 0.35, 0.5, rec2, 20, 40, 12, d["filtered2"], 1000, d["angles12"], d["xshift12"])
    err2 = np.abs(rec2 - d["recon2"]).max() / np.abs(d["recon2"]).max()
    print(f"backprojection vs reference: {err:.3e} {err2:.3e}")
    assert err2 < 1e-5


def test_fdk_chain_c1_water(torch):
    # forward c1 on the GPU, then filter + backproject on the GPU vs the oracle FDK of the same projections
    from virtualcbct_amd import fdk, io
    from virtualcbct_amd.projector import Projector
    import tempfile
    w = phantoms.WORKLOADS["c1"]
    with tempfile.TemporaryDirectory() as td:
        phantoms.write_xraydb(f"{td}/db")
        phantoms.write_spectrum_csv(f"{td}/s.csv")
        atten, pdf = io.attenuation_coefficient(f"{td}/s.csv", f"{td}/db")
        fm = f"{td}/m.flexmap"
        phantoms.write_flexmap(fm, 36, w["xshift_fdk"])
        angles_fdk, xs_fdk, ys_fdk = io.read_fdk_angles(fm, 36)
    vol = phantoms.water_cylinder(w["N"], w["Nz"], w["s"], w["radius"])
    ang, xs, ys = phantoms.flexmap_angles(36, w["xshift_proj"])
    cs, sn, xc, yc = io.trig_table(w["W"], w["H"], ang, xs, ys)
    P = Projector(w["N"], w["s"], w["Nz"], w["t"], w["H"], w["W"], 36, cs, sn, xc, yc)
    P.set_volume(vol)
    P.set_spectrum(atten, pdf)
    proj = P.forward(1000.0)
    filt = fdk.filter_sinogram_device(proj, xs_fdk, ys_fdk)
    rec = torch.zeros((w["Z"], w["R"], w["R"]), device="cuda")
    fdk.reconstruct_image_device(rec, filt, w["zstart"], w["zstep"], w["rscale"], 1000, angles_fdk, xs_fdk)
    rec = rec.cpu().numpy()
    proj_h = proj.cpu().numpy()
    filt_o = O.filter_sinogram(proj_h, xs_fdk, ys_fdk)
    zs = [0, 31, 63]
    rec_o = O.reconstruct_image(w["R"], w["Z"], w["zstart"], w["zstep"], w["rscale"], filt_o, 1000, angles_fdk,
                                xs_fdk)[zs]
    water = np.median(rec_o[1][120:150, 120:150])
    err = np.abs(rec[zs] - rec_o).max()
    print(f"c1 FDK: water level {water:.2f}, max |GPU - oracle| = {err:.4f} (tolerance {1e-3 * water:.3f})")
    assert 100 < water < 400
    assert err <= 1e-3 * water
    short = fdk.to_short(rec)
    assert short.dtype == np.int16


@pytest.mark.parametrize("R,cls", [(17, "M"), (270, "L"), (64, "S")])
def test_angle_mask_matches_oracle(torch, R, cls):
    """prepareAngles (F/misc.cpp:87-180) on the device, bit-exact against the oracle."""
    from virtualcbct_amd import fdk
    rng = np.random.default_rng(R)
    ang = np.concatenate([np.sort(rng.uniform(-np.pi, np.pi, 50)), [-np.pi, np.pi, 0.0]])
    got = fdk.prepareAngles(ang, R, 0.1, cls).cpu().numpy()
    assert np.array_equal(got, O.angle_mask(ang, R, 0.1, cls))


@pytest.mark.parametrize("taper", [0.0, 30.0])
def test_offset_weight_matches_oracle(torch, taper):
    """Offset-detector weighting and short-side padding on the device against the oracle (FP64 weights and
    powers on both sides; the FP32 results agree to an ulp or two), for positive, negative and zero shifts."""
    from virtualcbct_amd import fdk
    rng = np.random.default_rng(11)
    n, H, W = 5, 37, 400
    p = np.exp(-rng.uniform(0.0, 5.0, (n, H, W))).astype(np.float32)
    p[0, 0, :8] = [0.0, 1.0, 1e-30, 0.5, 1.0, 2.0, 0.25, 0.75]
    ys = np.array([0.0, 0.3, -1.0, 2.0, 0.1])
    d = torch.from_numpy(p).cuda()
    for xs in (np.array([11.5, 3.0, 0.0, 2.05, 11.43]), np.array([-11.5, -3.0, 0.0, -2.05, -11.43]), np.zeros(n)):
        want, xo_w, yo_w = O.offset_weight(p, xs, ys, taper)
        got, xo, yo, pit = fdk.offset_weight_device(d, xs, ys, taper)
        assert got.shape == want.shape
        np.testing.assert_allclose(got.cpu().numpy(), want, rtol=1e-6, atol=0)
        np.testing.assert_array_equal(xo, xo_w)
        np.testing.assert_array_equal(yo, yo_w)
        assert pit == pytest.approx(0.032509752 * want.shape[2] / W, rel=1e-15)
    with pytest.raises(Exception):
        fdk.offset_weight_device(d, np.array([20.0, 0, 0, 0, 0]), ys)
    with pytest.raises(Exception):
        fdk.offset_weight_device(d, np.array([3.0, -3.0, 0, 0, 0]), ys)


def test_offset_weight_half_fan_matches_centred_detector(torch):
    """c2's half-fan geometry (panel 115 pixels off centre, 85-pixel overlap) against a centred panel of the
    same width, which covers the whole 25 cm water cylinder (field radius 13 cm), on the device at c2's voxel
    size: with the weighting and padding, the half-fan reconstruction has the centred one's level at every
    radius (same spectrum, so the same beam-hardening cupping); plain FDK of the half-fan scan (the
    reference's behaviour) does not."""
    from virtualcbct_amd import fdk, io
    from virtualcbct_amd.projector import Projector
    import tempfile
    w = dict(phantoms.WORKLOADS["c2"])
    n_ang, Nz, Z, zstart = 120, 32, 10, -1
    R = w["R"]
    yy, xx = np.mgrid[0:R, 0:R]
    r = np.hypot(xx - (R - 1) / 2, yy - (R - 1) / 2) * w["rscale"]  # cm from the axis
    disk = r < 11.0
    vol = phantoms.water_cylinder(w["N"], Nz, w["s"], w["radius"])
    slices = {}
    with tempfile.TemporaryDirectory() as td:
        phantoms.write_xraydb(f"{td}/db")
        phantoms.write_spectrum_csv(f"{td}/s.csv")
        atten, pdf = io.attenuation_coefficient(f"{td}/s.csv", f"{td}/db")
        for name, xs_p, xs_f in (("centred", 0.0, 0.0), ("half-fan", w["xshift_proj"], w["xshift_fdk"])):
            fm = f"{td}/{name}.flexmap"
            phantoms.write_flexmap(fm, n_ang, xs_f)
            angles_fdk, xs_fdk, ys_fdk = io.read_fdk_angles(fm, n_ang)
            ang, xs, ys = phantoms.flexmap_angles(n_ang, xs_p)
            cs, sn, xc, yc = io.trig_table(w["W"], w["H"], ang, xs, ys)
            P = Projector(w["N"], w["s"], Nz, w["t"], w["H"], w["W"], n_ang, cs, sn, xc, yc)
            P.set_volume(vol)
            P.set_spectrum(atten, pdf)
            proj = P.forward(1000.0)
            P.close()
            rec = fdk.fdk_device(proj, zstart, w["zstep"], w["rscale"], R, Z, 1000, angles_fdk, xs_fdk, ys_fdk)
            slices[name] = rec[Z // 2].cpu().numpy().astype(np.float64)
            if name == "half-fan":
                rec = fdk.fdk_half_fan_device(proj, zstart, w["zstep"], w["rscale"], R, Z, 1000, angles_fdk,
                                              xs_fdk, ys_fdk)
                slices["half-fan weighted"] = rec[Z // 2].cpu().numpy().astype(np.float64)

    def rings(img):
        return np.array([img[(r >= a) & (r < a + 1)].mean() for a in range(11)])

    ref = slices["centred"]
    level = float(np.mean(ref[disk]))
    ring_w = float(np.abs(rings(slices["half-fan weighted"]) - rings(ref)).max()) / level
    ring_u = float(np.abs(rings(slices["half-fan"]) - rings(ref)).max()) / level
    pix_w = float(np.mean(np.abs(slices["half-fan weighted"] - ref)[disk])) / level
    print(f"half-fan vs centred panel (level {level:.1f} inside r < 11 cm): largest ring-mean difference "
          f"weighted {ring_w:.4f}, plain {ring_u:.4f}; mean |pixel difference| weighted {pix_w:.4f}")
    assert ring_w < 0.02
    assert pix_w < 0.04  # streaks of 120 views differ between the two panel positions
    assert ring_u > 0.2


def test_plan_ramlak_window_matches_oracle(torch):
    """The FDK plan's opt-in window alpha = 1 (band-limited ramp without the reference's window) against the
    oracle's filter with the same window (FP64), at the c2 panel width; the default plan stays the reference's
    kernel, and a wrong window parameter is refused."""
    from virtualcbct_amd import fdk
    rng = np.random.default_rng(7)
    n, H, W = 3, 12, 400
    proj = rng.uniform(0.05, 1.0, (n, H, W)).astype(np.float32)
    xs, ys = np.array([0.0, 11.5, -2.0]), np.array([0.0, 0.4, 1.0])
    ang = np.deg2rad(np.array([-170.0, -50.0, 70.0]))
    d = torch.from_numpy(proj).cuda()
    for alpha in (1.0, 0.5):
        plan = fdk.FdkPlan(H, W, n, 16, 2, -1, 0.1, 0.1, 1000, ang, xs, ys, window_alpha=alpha)
        out = plan.filter(d).cpu().numpy()
        plan.close()
        ref = O.filter_sinogram(proj, xs, ys, acc64=True, alpha=alpha)
        err = float(np.abs(out - ref).max() / np.abs(ref).max())
        print(f"plan filter, window alpha {alpha}: max err / max|ref| {err:.3e}")
        assert err < 2e-5
    with pytest.raises(Exception):
        fdk.FdkPlan(H, W, n, 16, 2, -1, 0.1, 0.1, 1000, ang, xs, ys, window_alpha=1.5)
