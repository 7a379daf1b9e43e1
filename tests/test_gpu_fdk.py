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


@pytest.mark.parametrize("W,nrows", [(33, 7), (700, 9), (1100, 5)])
def test_filter_other_widths_match_fp64_oracle(torch, W, nrows):
    """FFT sizes 128 (radix-2 last stage), 2048 (tables staged in LDS at the 64 KB limit) and 4096 (tables read
    from global memory), odd row counts (the last row pair has one row) and block tails."""
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
