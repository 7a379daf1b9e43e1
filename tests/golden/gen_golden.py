"""Writes the golden fixtures of tests/golden/ from the reference's own CPU code, compiled from its sources by
oracle/Makefile into oracle/_ref/ (run in the build container, where the reference tree exists):

    make -C oracle && python tests/golden/gen_golden.py

Every fixture holds seeded inputs (or the recipe to regenerate them with virtualcbct_amd.phantoms) and the
reference's outputs for them.
"""
import hashlib
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from virtualcbct_amd import phantoms  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"
P = f"{REF}/Release_CBCT_Projection_CM"
F = f"{REF}/Release_recon_from_forward_pro_float_CM"
R = f"{REF}/ICRPphantoms_2mmslice"


def small_proj_case():
    """32^2 x 8 random six-material phantom, 24 x 24 detector, 6 angles, 5 bins (pitch, geometry as physParams)."""
    rng = np.random.default_rng(20240601)
    N, Nz, H, W = 32, 8, 24, 24
    vol = rng.integers(0, 2000, (Nz, 6, N, N)).astype(np.uint16)
    ang = np.deg2rad(np.array([-170.0001, -90.00001, -40.00001, 9.99999, 59.99999, 133.3]))
    xs = np.full(6, 0.0123)
    ys = np.full(6, -0.007)
    NE = 5
    atten = rng.uniform(0.15, 0.5, (6, NE)).astype(np.float32)
    pdf = rng.uniform(0.1, 1, NE).astype(np.float32)
    pdf /= pdf.sum()
    return dict(N=N, Nz=Nz, H=H, W=W, s=0.5, t=0.5, z_reg=0.3, vol=vol, angles=ang, xshift=xs, yshift=ys,
                atten=atten, pdf=pdf, xfactor=1000.0)


def material_maps_case(seed=4242):
    """Label volume for the converter fixture: synthetic_labels(543, 271, 100, seed) with row 0 of slice 0
    set to 0..255 (then 0)."""
    lab = phantoms.synthetic_labels(543, 271, 100, seed=seed)
    lab[0, 0, :] = 0
    lab[0, 0, :256] = np.arange(256, dtype=np.uint8)
    return dict(seed=seed, labels=lab)


def iterative_fixture():
    """7. lengthproj and IR_ImageCBCT (P/lengthproj.cpp, P/IR_ImageCBCT.cpp) on a small seeded case."""
    rng = np.random.default_rng(31337)
    N, Nz, H, W, s, t = 20, 5, 24, 30, 0.6, 0.45
    ang = np.deg2rad(np.array([-150.0001, -60.5, 0.0, 45.0, 100.25, 170.0]))
    xs, ys = np.full(6, 0.021), np.full(6, -0.013)
    cs, sn, xc, yc = O.ref_trig_table(W, H, ang, xs, ys)
    g = O.Geometry(N=N, s=s, Nz=Nz, t=t, H=H, W=W, cos_t=cs, sin_t=sn, xc=xc, yc=yc, x_reg=0.2, z_reg=-0.3)
    proj = rng.integers(0, 65536, (6, H, W)).astype(np.uint16)
    npf = O.ref_lengthproj(g, proj, ang, 0)
    npf1 = O.ref_lengthproj(g, proj, ang, 1)
    before = rng.integers(0, 3000, (Nz, N, N)).astype(np.int16)
    npf_re = (npf * rng.uniform(0.8, 1.2, npf.shape)).astype(np.float32)
    ir = O.ref_ir_update(before, npf, npf_re, s, 0.3, 0.05)
    np.savez_compressed(f"{OUT}/iterative_small.npz", N=N, Nz=Nz, H=H, W=W, s=s, t=t, x_reg=0.2, z_reg=-0.3, angles=ang,
                        cos=cs, sin=sn, xc=xc, yc=yc, proj=proj, npf=npf, npf_ite1=npf1, before=before, npf_re=npf_re,
                        ir=ir)


def main():
    if not O.ref_available() or not O.ref_to_material_available():
        sys.exit("oracle/_ref not built (make -C oracle with the reference tree present)")

    # 1. attenuation_coefficient on the shipped spectrum and xraydb tables
    at, pdf = O.ref_attenuation_coefficient(f"{P}/Spectrum/result_20220111_average.csv", P)
    np.savez_compressed(f"{OUT}/atten_shipped.npz", atten=at, pdf=pdf)

    # 2. attenuation_coefficient on the synthetic 100 kV spectrum and synthetic xraydb (regenerable)
    with tempfile.TemporaryDirectory() as td:
        os.makedirs(f"{td}/attenuation_xraydb")
        phantoms.write_xraydb(f"{td}/attenuation_xraydb", seed=1234)
        phantoms.write_spectrum_csv(f"{td}/spec100.csv", kvp=100, e0=10)
        phantoms.write_spectrum_csv(f"{td}/spec80.csv", kvp=80, e0=10)
        a1, p1 = O.ref_attenuation_coefficient(f"{td}/spec100.csv", td)
        a2, p2 = O.ref_attenuation_coefficient(f"{td}/spec80.csv", td)
        a3, p3 = O.ref_attenuation_coefficient(f"{td}/spec100.csv", td, nthin=3)
    np.savez_compressed(f"{OUT}/atten_synth.npz", atten100=a1, pdf100=p1, atten80=a2, pdf80=p2, atten100_thin3=a3,
                        pdf100_thin3=p3, xraydb_seed=1234)

    # 3. FDK pieces on small inputs
    rng = np.random.default_rng(777)
    proj = rng.uniform(0.05, 1.0, (4, 8, 16)).astype(np.float32)
    xs4 = rng.uniform(-1, 1, 4)
    ys4 = rng.uniform(-1, 1, 4)
    cm16 = O.ref_const_map(16)
    cm400 = O.ref_const_map(400)
    filt, _ = O.ref_filter_sinogram(proj, xs4, ys4)
    sino = rng.normal(0, 1, (8, 8, 16)).astype(np.float32)
    ang8 = np.sort(rng.uniform(-np.pi, np.pi, 8))
    xs8 = rng.uniform(-0.5, 0.5, 8)
    rec, _ = O.ref_reconstruct_image(16, 4, -1, 0.5, 0.3, sino, 1000, ang8, xs8)
    # a wider case through the whole filter + backprojection chain
    proj2 = rng.uniform(0.2, 1.0, (12, 20, 40)).astype(np.float32)
    ang12 = np.deg2rad(np.arange(12) * 30.0 - 170.0001)
    xs12 = np.full(12, 0.3)
    ys12 = np.zeros(12)
    filt2, _ = O.ref_filter_sinogram(proj2, xs12, ys12)
    rec2, _ = O.ref_reconstruct_image(24, 6, -1, 0.35, 0.5, filt2, 1000, ang12, xs12)
    np.savez_compressed(f"{OUT}/fdk_small.npz", proj=proj, xshift=xs4, yshift=ys4, const_map16=cm16,
                        const_map400_row0=cm400[0], const_map400_row199=cm400[199], filtered=filt, sino=sino,
                        angles=ang8, xshift8=xs8, recon=rec, mask16=O.ref_mask(16), mask270=O.ref_mask(270),
                        proj2=proj2, angles12=ang12, xshift12=xs12, yshift12=ys12, filtered2=filt2, recon2=rec2)

    # 4. the CPU twin of the projector on a small random phantom, every (angle, bin)
    c = small_proj_case()
    cs, sn, xc, yc = O.ref_trig_table(c["W"], c["H"], c["angles"], c["xshift"], c["yshift"])
    g = O.Geometry(N=c["N"], s=c["s"], Nz=c["Nz"], t=c["t"], H=c["H"], W=c["W"], cos_t=cs, sin_t=sn, xc=xc, yc=yc,
                   z_reg=c["z_reg"])
    twin = np.zeros((len(c["angles"]), c["atten"].shape[1], c["H"], c["W"]), np.float32)
    for a in range(len(c["angles"])):
        for ke in range(c["atten"].shape[1]):
            twin[a, ke] = O.ref_reprojection_one(g, c["vol"], c["angles"], c["xfactor"], c["atten"], c["pdf"], a, ke)
    np.savez_compressed(f"{OUT}/reproj_small.npz", twin=twin, cos=cs, sin=sn, xc=xc, yc=yc,
                        **{k: v for k, v in c.items() if isinstance(v, np.ndarray)})

    # 5. loaders: FDK angle lists (CW/CCW reversal) and the projector's flexmap reading, shipped files
    out = {}
    for name, n in (("36_little_slide.flexmap", 36), ("360_littele_slide.flexmap", 360)):
        txt = f"{F}/txtfile/FDK_{'36pro_water20cm' if n == 36 else '360pro_water25cm'}.txt"
        d = O.ref_fdk_load(txt, F)
        key = f"fdk{n}"
        out[key + "_angles"] = d["angles"]
        out[key + "_ints"] = d["ints"]
        out[key + "_dbls"] = d["dbls"]
    for n, txt in ((36, "IR_waterphantom_36pro.txt"), (360, "IR_waterphantom_360pro.txt")):
        d = O.ref_proj_load(f"{P}/txtfile/{txt}", P)
        out[f"proj{n}_angles"] = d["angles"]
        out[f"proj{n}_ints"] = d["ints"]
        out[f"proj{n}_dbls"] = d["dbls"]
    np.savez_compressed(f"{OUT}/loaders.npz", **out)

    # 6. the phantom converter (R = ICRPphantoms_2mmslice) on a seeded synthetic label volume of its fixed
    #    shape; row 0 of slice 0 holds every label 0..255, so the reference's output there is its whole
    #    label -> map table.  Kept: that table, the printed media weights, and digests of the outputs.
    md_fixture = material_maps_case()
    with tempfile.TemporaryDirectory() as td:
        whole, md, stdout = O.ref_to_material_run(md_fixture["labels"], f"{R}/AM_media.csv", f"{R}/AM_organs.csv", td)
    lines = stdout.splitlines()
    i0 = lines.index("AM_media.csv ") + 1
    media = np.array([[float(v) for v in ln.split()] for ln in lines[i0:i0 + 53]])
    np.savez_compressed(f"{OUT}/material_maps.npz", seed=md_fixture["seed"], lut=md[0, :, 0, :256].T.copy(),
                        media_printed=media, md_sha256=np.bytes_(hashlib.sha256(md.tobytes()).hexdigest()),
                        whole_sha256=np.bytes_(hashlib.sha256(whole.tobytes()).hexdigest()),
                        md_sums=md.astype(np.int64).sum(axis=(2, 3)))
    iterative_fixture()
    print("golden fixtures written to", OUT)


if __name__ == "__main__":
    main()
