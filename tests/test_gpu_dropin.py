"""The drop-in programs on reference-format inputs: the reference's own P/main.cpp and F/main.cpp linked
against libvcbct_compat.so (oracle/_ref/proj.exe, oracle/_ref/cbct2, built by `make -C oracle dropin`) and
this build's host programs (virtualcbct_amd/bin/vcbct_proj, vcbct_fdk), compared with each other and with the
FP64 oracle."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from virtualcbct_amd import io, phantoms

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_PROJ = os.path.join(ROOT, "oracle", "_ref", "proj.exe")
REF_FDK = os.path.join(ROOT, "oracle", "_ref", "cbct2")
MY_PROJ = os.path.join(ROOT, "virtualcbct_amd", "bin", "vcbct_proj")
MY_FDK = os.path.join(ROOT, "virtualcbct_amd", "bin", "vcbct_fdk")

N, NZ, S, T, H, W, NA = 64, 8, 0.25, 0.5, 80, 80, 36


def setup_dir(d):
    os.makedirs(d, exist_ok=True)
    phantoms.write_xraydb(os.path.join(d, "attenuation_xraydb"))
    phantoms.write_spectrum_csv(os.path.join(d, "spec.csv"), kvp=100)
    phantoms.write_flexmap(os.path.join(d, "m.flexmap"), NA)
    vol = phantoms.water_cylinder(N, NZ, S, 6.0)
    vol[:, 4, 20:30, 30:40] = 150  # a phosphorus insert to make the materials matter
    phantoms.write_slices(vol, os.path.join(d, "ph", "w"))
    phantoms.write_proj_txt(os.path.join(d, "p.txt"), "spec.csv", NA, W, H, "m.flexmap", "ph/w", NZ // 2, "proj",
                            (S, S, T), N, NZ)
    phantoms.write_fdk_txt(os.path.join(d, "f.txt"), NA, W, H, "proj", "raw", "m.flexmap", 48, -2, 0.5, 8, 0.3, 1000,
                           "rec_short.raw")
    return vol


def run(exe, d, txt):
    r = subprocess.run([exe, txt], cwd=d, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]


@pytest.fixture(scope="module")
def torch():
    import torch as T
    if not T.cuda.is_available():
        pytest.skip("no GPU")
    return T


def test_dropin_projector_and_fdk(torch, tmp_path):
    if not (os.path.exists(MY_PROJ) and os.path.exists(MY_FDK)):
        pytest.skip("host programs not built")
    d1 = str(tmp_path / "mine")
    vol = setup_dir(d1)
    run(MY_PROJ, d1, "p.txt")
    mine = np.fromfile(os.path.join(d1, "proj.raw"), np.float32).reshape(NA, H, W)
    # oracle of the same inputs
    atten, pdf = io.attenuation_coefficient(os.path.join(d1, "spec.csv"), os.path.join(d1, "attenuation_xraydb"))
    ang, xs, ys = io.read_flexmap(os.path.join(d1, "m.flexmap"))
    cs, sn, xc, yc = io.trig_table(W, H, ang, xs, ys)
    g = O.Geometry(N=N, s=S, Nz=NZ, t=T, H=H, W=W, cos_t=cs, sin_t=sn, xc=xc, yc=yc)
    ref = O.forward_collapsed(g, vol, atten, pdf, 1000.0)
    e = float(np.max(np.abs(mine - ref) / ref))
    print(f"vcbct_proj vs FP64 oracle: {e:.3e}")
    assert e < 1e-4
    run(MY_FDK, d1, "f.txt")
    rec_mine = np.fromfile(os.path.join(d1, "rec_short.raw"), np.int16).reshape(8, 48, 48)
    a_f, xs_f, ys_f = io.read_fdk_angles(os.path.join(d1, "m.flexmap"), NA)
    filt = O.filter_sinogram(mine, xs_f, ys_f)
    rec_o = O.reconstruct_image(48, 8, -2, 0.5, 0.3, filt, 1000, a_f, xs_f)
    assert np.abs(rec_mine.astype(np.int32) - np.trunc(rec_o).astype(np.int32)).max() <= 1
    flt = np.fromfile(os.path.join(d1, "fltsngrm.raw"), np.float32).reshape(NA, H, W)
    assert np.abs(flt - filt).max() <= 2e-5 * np.abs(filt).max()

    if not (os.path.exists(REF_PROJ) and os.path.exists(REF_FDK)):
        pytest.skip("reference host programs (oracle/_ref) not built")
    d2 = str(tmp_path / "ref")
    setup_dir(d2)
    run(REF_PROJ, d2, "p.txt")
    theirs = np.fromfile(os.path.join(d2, "proj.raw"), np.float32).reshape(NA, H, W)
    assert np.array_equal(theirs, mine)
    run(REF_FDK, d2, "f.txt")
    rec_theirs = np.fromfile(os.path.join(d2, "rec_short.raw"), np.int16).reshape(8, 48, 48)
    assert np.array_equal(rec_theirs, rec_mine)


def test_dropin_noise_and_half_fan_options(torch, tmp_path):
    """The host programs' opt-in stages (not in the reference): VCBCT_NOISE_I0 / VCBCT_NOISE_SEED make
    vcbct_proj write the detector model's counts of its plain output, equal to the library call on the device;
    VCBCT_HALF_FAN=1 makes vcbct_fdk weight and pad an offset panel, equal to fdk_half_fan_device before the
    int16 truncation (both use the one-shot filter and backprojector)."""
    if not (os.path.exists(MY_PROJ) and os.path.exists(MY_FDK)):
        pytest.skip("host programs not built")
    from virtualcbct_amd import detector, fdk
    d = str(tmp_path / "opt")
    setup_dir(d)
    run(MY_PROJ, d, "p.txt")
    plain = np.fromfile(os.path.join(d, "proj.raw"), np.float32).reshape(NA, H, W)
    env = dict(os.environ, VCBCT_NOISE_I0="20000", VCBCT_NOISE_SEED="7")
    r = subprocess.run([MY_PROJ, "p.txt"], cwd=d, env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    noisy = np.fromfile(os.path.join(d, "proj.raw"), np.float32).reshape(NA, H, W)
    want = detector.detector_counts(torch.from_numpy(plain).cuda(), I0=20000.0, seed=7).cpu().numpy()
    assert np.array_equal(noisy, want)
    assert not np.array_equal(noisy, plain)
    # the FDK program on an offset panel (12 of the 40 half-panel pixels: 6.0 in the FDK's units, 0.5 per pixel)
    phantoms.write_flexmap(os.path.join(d, "m.flexmap"), NA, 6.0)
    env = dict(os.environ, VCBCT_HALF_FAN="1")
    r = subprocess.run([MY_FDK, "f.txt"], cwd=d, env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    rec_prog = np.fromfile(os.path.join(d, "rec_short.raw"), np.int16).reshape(8, 48, 48)
    a_f, xs_f, ys_f = io.read_fdk_angles(os.path.join(d, "m.flexmap"), NA)
    rec = fdk.fdk_half_fan_device(torch.from_numpy(noisy).cuda(), -2, 0.5, 0.3, 48, 8, 1000, a_f, xs_f, ys_f)
    want_rec = np.trunc(rec.cpu().numpy().astype(np.float64)).astype(np.int32)
    # the program accumulates in FP64 on the host path, the device path in FP32: allow one unit of truncation
    assert np.abs(rec_prog.astype(np.int32) - want_rec).max() <= 1
    assert np.abs(rec_prog).max() > 10
    flt = np.fromfile(os.path.join(d, "fltsngrm.raw"), np.float32)
    assert flt.size == NA * H * fdk.half_fan_width(W, xs_f)
