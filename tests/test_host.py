"""CPU tests of the product library's host side: it loads and exports every entry point of include/vcbct.h,
its host formats agree with the reference's loaders (golden fixtures) and with the oracle, and its error
paths return codes instead of exiting.  No kernel is launched here."""
import ctypes as C
import os
import re
import tempfile

import numpy as np
import pytest

from oracle import oracle as O
from virtualcbct_amd import _lib, io, phantoms
from virtualcbct_amd._lib import VcbctError

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def header_symbols(name):
    txt = open(os.path.join(ROOT, "include", name)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(vcbct_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    syms = header_symbols("vcbct.h")
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) <= set(_lib.SIGNATURES), set(syms) - set(_lib.SIGNATURES)
    assert b"gfx950" in L.vcbct_version()


def test_compat_library_exports_reference_entry_points():
    path = os.path.join(ROOT, "virtualcbct_amd", "libvcbct_compat.so")
    if not os.path.exists(path):
        pytest.skip("compat shim not built")
    Lc = C.CDLL(path)
    # C++-linkage names of initializeGPU / reprojectionOnGPU / terminateGPU (P/reprojectionOnGPU.h:2-19) and
    # prepareConstMap / filterSinogram / reconstructImage (F/filterSinogram.h, F/reconstructImage.h)
    for mangled in ("_Z13initializeGPUididiiiPdS_S_S_dddddddi", "_Z12terminateGPUv",
                    "_Z17reprojectionOnGPUiiiPtPfS0_S0_ddiiiiPdS1_S1_S1_S1_dS1_dddddddS0_S0_iiPc",
                    "_Z15prepareConstMapid", "_Z14filterSinogramiiiPPPfS1_PPdS2_S2_d",
                    "_Z16reconstructImageiiiddPPPdPPciiiPPPfiS_S_S_S_d"):
        assert hasattr(Lc, mangled), mangled


def test_const_map_matches_reference():
    d = np.load(os.path.join(GOLD, "fdk_small.npz"))
    from virtualcbct_amd import fdk
    assert np.array_equal(fdk.prepareConstMap(16), d["const_map16"])
    cm = fdk.prepareConstMap(400)
    assert np.array_equal(cm[0], d["const_map400_row0"])
    assert np.array_equal(cm[199], d["const_map400_row199"])


@pytest.fixture(scope="module")
def synth_db(tmp_path_factory):
    d = tmp_path_factory.mktemp("db")
    phantoms.write_xraydb(str(d / "attenuation_xraydb"), seed=1234)
    phantoms.write_spectrum_csv(str(d / "spec100.csv"), kvp=100, e0=10)
    phantoms.write_spectrum_csv(str(d / "spec80.csv"), kvp=80, e0=10)
    return d


def test_attenuation_table_matches_reference_synthetic(synth_db):
    g = np.load(os.path.join(GOLD, "atten_synth.npz"))
    db = str(synth_db / "attenuation_xraydb")
    for kv in (100, 80):
        at, pdf = io.attenuation_coefficient(str(synth_db / f"spec{kv}.csv"), db)
        assert np.array_equal(at, g[f"atten{kv}"]) and np.array_equal(pdf, g[f"pdf{kv}"])
        at2, pdf2 = O.attenuation_table(str(synth_db / f"spec{kv}.csv"), db)
        assert np.array_equal(at2, g[f"atten{kv}"]) and np.array_equal(pdf2, g[f"pdf{kv}"])
    at, pdf = io.attenuation_coefficient(str(synth_db / "spec100.csv"), db, nthin=3)
    assert np.array_equal(at, g["atten100_thin3"]) and np.array_equal(pdf, g["pdf100_thin3"])


def test_attenuation_table_matches_reference_shipped(ref_dir):
    g = np.load(os.path.join(GOLD, "atten_shipped.npz"))
    P = os.path.join(ref_dir, "Release_CBCT_Projection_CM")
    at, pdf = io.attenuation_coefficient(f"{P}/Spectrum/result_20220111_average.csv", f"{P}/attenuation_xraydb")
    assert at.shape == (6, 101)
    assert np.array_equal(at, g["atten"]) and np.array_equal(pdf, g["pdf"])


def test_trig_table_matches_reference():
    d = np.load(os.path.join(GOLD, "reproj_small.npz"))
    cs, sn, xc, yc = io.trig_table(24, 24, d["angles"], d["xshift"], d["yshift"])
    for a, b in ((cs, "cos"), (sn, "sin"), (xc, "xc"), (yc, "yc")):
        assert np.array_equal(a, d[b])


def test_flexmaps_and_angle_loaders(tmp_path):
    d = np.load(os.path.join(GOLD, "loaders.npz"))
    for n in (36, 360):
        fm = str(tmp_path / f"m{n}.flexmap")
        phantoms.write_flexmap(fm, n)
        a, _, _ = io.read_fdk_angles(fm, n)
        assert np.array_equal(a, d[f"fdk{n}_angles"])
        a2, x2, y2 = io.read_flexmap(fm)
        assert np.array_equal(a2, d[f"proj{n}_angles"])
        a3, _, _ = phantoms.flexmap_angles(n)
        assert np.array_equal(a3, d[f"proj{n}_angles"])


def test_generators_reproduce_shipped_files(ref_dir):
    P = os.path.join(ref_dir, "Release_CBCT_Projection_CM")
    for n, name in ((36, "36_little_slide.flexmap"), (360, "360_little_slide.flexmap")):
        ref_rows = [r.split() for r in open(os.path.join(P, name)).read().split("\n") if r.strip()]
        mine = [r.split() for r in phantoms.flexmap_rows(n)]
        assert [float(r[0]) for r in ref_rows] == [float(r[0]) for r in mine]
    for r, folder, stem in ((10.0, "waterphan_20cm", "Catphan503_water20cm_Weight_input_512x512x6_short"),
                            (12.5, "waterphan_25cm", "waterphantom_Weight_input_512x512x6_short")):
        shipped = np.fromfile(os.path.join(ref_dir, "waterphantom", folder, f"{stem}_000.raw"), np.uint16)
        mine = phantoms.water_cylinder(512, 1, 0.1, r)[0]
        assert np.array_equal(mine.ravel(), shipped)


def test_config_readers_match_reference_loaders(ref_dir):
    d = np.load(os.path.join(GOLD, "loaders.npz"))
    P = os.path.join(ref_dir, "Release_CBCT_Projection_CM")
    F = os.path.join(ref_dir, "Release_recon_from_forward_pro_float_CM")
    for n, txt in ((36, "IR_waterphantom_36pro.txt"), (360, "IR_waterphantom_360pro.txt")):
        c = io.read_proj_config(os.path.join(P, "txtfile", txt))
        ints, dbls = d[f"proj{n}_ints"], d[f"proj{n}_dbls"]
        assert (c["total_proj"], c["proj_width"], c["proj_height"], c["recon_size"], c["mat_num"]) == \
            (ints[0], ints[1], ints[2], ints[4], ints[6])
        lo, hi = io.slice_window(c)
        assert hi - lo == ints[7] and c["zic"] - lo == ints[3]
        assert (c["x_scale"], c["y_scale"], c["z_scale"], c["xfactor"]) == tuple(dbls[:4])
    for n, txt in ((36, "FDK_36pro_water20cm.txt"), (360, "FDK_360pro_water25cm.txt")):
        c = io.read_fdk_config(os.path.join(F, "txtfile", txt))
        ints, dbls = d[f"fdk{n}_ints"], d[f"fdk{n}_dbls"]
        assert (c["total_proj"], c["proj_width"], c["proj_height"], c["recon_size"], c["recon_offset_cm"],
                c["recon_slices"], c["xfactor"]) == tuple(ints[:7])
        assert (c["recon_step"], c["recon_scale"]) == tuple(dbls[:2])


def test_txt_writers_roundtrip(tmp_path):
    t = str(tmp_path / "p.txt")
    phantoms.write_proj_txt(t, "spec.csv", 36, 400, 400, "m.flexmap", "ph/w", 32, "out", (0.05, 0.05, 0.1), 512, 64)
    c = io.read_proj_config(t)
    assert c["total_proj"] == 36 and c["recon_slices"] == 64 and c["z_scale"] == 0.1 and c["mat_prefix"] == "ph/w"
    lo, hi = io.slice_window(c)
    assert (lo, hi) == (0, 64)
    f = str(tmp_path / "f.txt")
    phantoms.write_fdk_txt(f, 36, 400, 400, "proj", "raw", "m.flexmap", 270, -3, 0.1, 64, 0.1, 1000, "o.raw")
    c = io.read_fdk_config(f)
    assert (c["recon_size"], c["recon_offset_cm"], c["recon_slices"], c["xfactor"], c["ext"]) == (270, -3, 64, 1000,
                                                                                                     "raw")


def test_material_slices_and_missing_policy(tmp_path):
    vol = phantoms.water_cylinder(16, 3, 1.0, 5.0)
    phantoms.write_slices(vol, str(tmp_path / "ph"))
    got, miss = io.load_material_slices(str(tmp_path / "ph"), 0, 3, 16, 6)
    assert np.array_equal(got, vol) and miss == 0
    with pytest.raises(VcbctError) as e:
        io.load_material_slices(str(tmp_path / "ph"), 0, 5, 16, 6)
    assert e.value.code == _lib.VCBCT_EIO
    got, miss = io.load_material_slices(str(tmp_path / "ph"), 0, 5, 16, 6, missing_policy=1)
    assert miss == 2 and np.array_equal(got[4], vol[2])  # the reference repeats the last slice read


def test_error_paths_return_codes(tmp_path):
    with pytest.raises(VcbctError) as e:
        io.read_proj_config(str(tmp_path / "missing.txt"))
    assert e.value.code == _lib.VCBCT_EIO
    with pytest.raises(VcbctError) as e:
        io.attenuation_coefficient(str(tmp_path / "nope.csv"), str(tmp_path))
    assert e.value.code == _lib.VCBCT_EIO
    # invalid geometry is rejected before any device is touched
    L = _lib.lib()
    g = _lib.vcbct_geometry(0, 0.05, 64, 0.1, 400, 400, 36, 0, 0, 0, 0.05, 0.0325, 100, 800, 6)
    h = C.c_void_p()
    z = np.zeros(36)
    p = z.ctypes.data_as(_lib.P_D)
    assert L.vcbct_projector_create(C.byref(h), C.byref(g), p, p, p, p, 0) == _lib.VCBCT_EINVAL
    g.recon_size, g.num_material = 512, 7
    assert L.vcbct_projector_create(C.byref(h), C.byref(g), p, p, p, p, 0) == _lib.VCBCT_EINVAL
    assert b"num_material" in L.vcbct_last_error()
    # malformed xraydb (truncated) is an error, not a silent repeat
    d = tmp_path / "db"
    d.mkdir()
    for nm in ("photo", "comp", "rayl"):
        (d / f"{nm}_5keVto150keV.csv").write_text("H,C,N,O,P,Ca,energy\n0.1,0.1,0.1,0.1,0.1,0.1,5000\n")
    phantoms.write_spectrum_csv(str(tmp_path / "s.csv"))
    with pytest.raises(VcbctError) as e:
        io.attenuation_coefficient(str(tmp_path / "s.csv"), str(d))
    assert e.value.code == _lib.VCBCT_EIO


def _build_shim_driver(tmp_path):
    """The compat shim's source (compat_shim.cpp) linked with a host stand-in of the four C-ABI calls it makes
    (tests/shim_host/stub_vcbct.cpp) and a driver that calls it as P/main.cpp:111-125 does."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    src = [os.path.join(ROOT, "virtualcbct_amd", "csrc", "compat_shim.cpp"),
           os.path.join(ROOT, "tests", "shim_host", "stub_vcbct.cpp"),
           os.path.join(ROOT, "tests", "shim_host", "drive_shim.cpp")]
    exe = str(tmp_path / "drive_shim")
    base = ["g++", "-std=c++17", "-O1", "-g", "-I", os.path.join(ROOT, "include"), *src, "-o", exe]
    # address sanitizer on the host build when the toolchain has it: an overrun of the output buffer fails loudly
    r = subprocess.run(base[:1] + ["-fsanitize=address", "-fno-omit-frame-pointer"] + base[1:], capture_output=True)
    if r.returncode != 0:
        subprocess.run(base, check=True)
    return exe


def test_compat_shim_reprojection_angle_range(tmp_path):
    """reprojectionOnGPU writes angles [0, endProjNumber) (P/reprojectionOnGPU.cu:976), sized by endProjNumber;
    an endProjNumber beyond initializeGPU's totalProjNumber, below 1, or an image size other than
    initializeGPU's exits with status 1 and a message, never writing past the buffer."""
    import subprocess
    exe = _build_shim_driver(tmp_path)
    H, W = 3, 5
    a = np.arange(4)[:, None, None] * 1e6 + np.arange(H)[None, :, None] * 1e3 + np.arange(W)[None, None, :]
    for total, end in ((4, 4), (4, 2), (4, 1)):
        out = tmp_path / f"proj_{total}_{end}"
        r = subprocess.run([exe, str(total), str(end), str(out)], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        got = np.fromfile(str(out) + ".raw", np.float32)
        assert got.size == end * H * W
        assert np.array_equal(got.reshape(end, H, W), a[:end].astype(np.float32))
    for total, end, extra, msg in ((4, 6, [], "6 angles requested"), (4, 0, [], "must be positive"),
                                   (4, 4, ["7", "3"], "differs from initializeGPU")):
        out = tmp_path / f"bad_{end}"
        r = subprocess.run([exe, str(total), str(end), str(out), *extra], capture_output=True, text=True)
        assert r.returncode == 1 and msg in r.stderr, (r.returncode, r.stderr)
        assert not os.path.exists(str(out) + ".raw")


def test_filter_fft_length_rule():
    """The filter's transform length for every panel width it takes: at least 2W - 1 (so the circular
    convolution equals the linear one on the W outputs), a power of two >= 64 or 5 x 2^m >= 80, and the shortest
    such length; out-of-range widths are errors, not lengths."""
    from virtualcbct_amd import fdk
    pow2 = [1 << k for k in range(6, 13)]
    five = [80 << k for k in range(0, 7)]
    for W in range(1, 2049):
        L = fdk.filter_fft_length(W)
        ok = [c for c in pow2 + five if c >= 2 * W - 1]
        assert L == min(ok), (W, L)
    assert [fdk.filter_fft_length(W) for W in (16, 33, 300, 400, 630, 700, 1100, 2048)] == \
        [64, 80, 640, 1024, 1280, 2048, 2560, 4096]
    for W in (0, -3, 2049):
        with pytest.raises(VcbctError):
            fdk.filter_fft_length(W)
