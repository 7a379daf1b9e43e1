"""Synthetic inputs in the reference's formats: water-cylinder and ellipsoid phantoms (uint16 elemental
density slices, g/cc x 1000, materials H, C, N, O, P, Ca), a Kramers-plus-tungsten-peaks spectrum CSV,
xraydb-format attenuation CSVs from a photoelectric Z^4/E^3 + Klein-Nishina model, flexmaps, and the
projector / FDK txt files.  All generators are seeded and deterministic.
"""
from __future__ import annotations

import os

import numpy as np

MATERIALS = ("H", "C", "N", "O", "P", "Ca")
Z_OF = np.array([1, 6, 7, 8, 15, 20], np.float64)
A_OF = np.array([1.008, 12.011, 14.007, 15.999, 30.974, 40.078])
R_E_CM = 2.8179403262e-13
N_A = 6.02214076e23


# --------------------------------------------------------------------------------------------- phantoms
def water_cylinder(N, Nz, scale_cm, radius_cm, water=(111, 888), air_n=1):
    """[z][mat][y][x] uint16.  Inside (i - N/2)^2 + (j - N/2)^2 <= (r/s)^2 (voxel units): H, O = water;
    elsewhere N = air_n.  At N=512, s=0.1 cm this reproduces waterphantom/waterphan_20cm (r = 10 cm) and
    waterphan_25cm (r = 12.5 cm) slice for slice."""
    i = np.arange(N, dtype=np.float64) - N / 2
    X, Y = np.meshgrid(i, i)
    R = radius_cm / scale_cm
    inside = (X * X + Y * Y) <= R * R
    sl = np.zeros((6, N, N), np.uint16)
    sl[0][inside] = water[0]
    sl[3][inside] = water[1]
    sl[2][~inside] = air_n
    return np.ascontiguousarray(np.broadcast_to(sl, (Nz, 6, N, N)))


def water_cylinder_f32(N, Nz, scale_cm, radius_cm):
    """FP32 maps [mat][z][y][x] in g/cc of the same cylinder (H 0.111, O 0.888, air N 0.0012)."""
    i = np.arange(N, dtype=np.float64) - N / 2
    X, Y = np.meshgrid(i, i)
    inside = (X * X + Y * Y) <= (radius_cm / scale_cm) ** 2
    m = np.zeros((6, N, N), np.float32)
    m[0][inside] = 0.111
    m[3][inside] = 0.888
    m[2][~inside] = 0.0012
    return np.ascontiguousarray(np.broadcast_to(m[:, None], (6, Nz, N, N)))


def ellipsoid_phantom(N, Nz, scale_cm, step_cm, seed, n_ell=40, n_labels=8):
    """Anthropomorphic-like phantom: union of `n_ell` seeded random ellipsoids with tissue labels.  Each label
    has a Dirichlet(1) mass-fraction vector over the six elements and a density U(0.9, 1.9) g/cc; 'bone' labels
    (the last two) have P and Ca boosted x10 before renormalisation.  Air (N 0.0012 g/cc) elsewhere.
    Returns uint16 [z][mat][y][x] (g/cc x 1000, truncated)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    frac = rng.dirichlet(np.ones(6), size=n_labels)
    frac[-2:, 4:] *= 10.0
    frac /= frac.sum(1, keepdims=True)
    dens = rng.uniform(0.9, 1.9, n_labels)
    label = np.full((Nz, N, N), -1, np.int16)
    zc = (np.arange(Nz) - Nz / 2 + 0.5) * step_cm
    yc = (np.arange(N) - N / 2 + 0.5) * scale_cm
    half_xy = N * scale_cm / 2
    half_z = Nz * step_cm / 2
    Zg, Yg, Xg = np.meshgrid(zc, yc, yc, indexing="ij")
    for k in range(n_ell):
        c = rng.uniform([-0.5 * half_xy, -0.5 * half_xy, -0.7 * half_z], [0.5 * half_xy, 0.5 * half_xy, 0.7 * half_z])
        ax = rng.uniform([0.1 * half_xy, 0.1 * half_xy, 0.1 * half_z], [0.45 * half_xy, 0.45 * half_xy, 0.6 * half_z])
        th = rng.uniform(0, np.pi)
        lab = rng.integers(0, n_labels)
        ct, st = np.cos(th), np.sin(th)
        dx, dy, dz = Xg - c[0], Yg - c[1], Zg - c[2]
        u = (ct * dx + st * dy) / ax[0]
        v = (-st * dx + ct * dy) / ax[1]
        w = dz / ax[2]
        label[(u * u + v * v + w * w) <= 1.0] = lab
    vol = np.zeros((Nz, 6, N, N), np.uint16)
    for lab in range(n_labels):
        msk = label == lab
        for m in range(6):
            vol[:, m][msk] = np.uint16(int(frac[lab, m] * dens[lab] * 1000.0))
    vol[:, 2][label < 0] = 1
    return vol


def ellipsoid_phantom_torch(N, Nz, scale_cm, step_cm, seed, n_ell=40, n_labels=8, device="cuda"):
    """The same phantom as ellipsoid_phantom (same seeded parameters, same label rule), voxelised on the device
    with torch for large volumes; returns an int16 device tensor [z][mat][y][x] holding the uint16 values."""
    import torch
    rng = np.random.Generator(np.random.PCG64(seed))
    frac = rng.dirichlet(np.ones(6), size=n_labels)
    frac[-2:, 4:] *= 10.0
    frac /= frac.sum(1, keepdims=True)
    dens = rng.uniform(0.9, 1.9, n_labels)
    half_xy = N * scale_cm / 2
    half_z = Nz * step_cm / 2
    zc = torch.tensor((np.arange(Nz) - Nz / 2 + 0.5) * step_cm, dtype=torch.float64, device=device)
    yc = torch.tensor((np.arange(N) - N / 2 + 0.5) * scale_cm, dtype=torch.float64, device=device)
    label = torch.full((Nz, N, N), -1, dtype=torch.int16, device=device)
    Zg, Yg, Xg = torch.meshgrid(zc, yc, yc, indexing="ij")
    for k in range(n_ell):
        c = rng.uniform([-0.5 * half_xy, -0.5 * half_xy, -0.7 * half_z], [0.5 * half_xy, 0.5 * half_xy, 0.7 * half_z])
        ax = rng.uniform([0.1 * half_xy, 0.1 * half_xy, 0.1 * half_z], [0.45 * half_xy, 0.45 * half_xy, 0.6 * half_z])
        th = rng.uniform(0, np.pi)
        lab = int(rng.integers(0, n_labels))
        ct, st = np.cos(th), np.sin(th)
        dx, dy, dz = Xg - c[0], Yg - c[1], Zg - c[2]
        u = (ct * dx + st * dy) / ax[0]
        v = (-st * dx + ct * dy) / ax[1]
        w = dz / ax[2]
        label[(u * u + v * v + w * w) <= 1.0] = lab
        del dx, dy, dz, u, v, w
    vals = np.zeros((n_labels + 1, 6), np.int32)
    for lab in range(n_labels):
        for m in range(6):
            vals[lab, m] = int(frac[lab, m] * dens[lab] * 1000.0)
    vals[n_labels, 2] = 1  # air (label -1 maps to the last row)
    lut = torch.tensor(vals, dtype=torch.int32, device=device)
    idx = torch.where(label < 0, torch.full_like(label, n_labels), label).long()
    vol = lut[idx]                       # [z][y][x][mat]
    return vol.permute(0, 3, 1, 2).contiguous().to(torch.int16)


def synthetic_labels(ncol=543, nrow=271, nz=100, seed=0, labels=None, n_organs=60):
    """Organ-label volume uint8 [nz][nrow][ncol] in the layout of the ICRP label slices the reference's
    converter reads (ICRPphantoms_2mmslice/To_material_distribution_for512_random.cpp:38, 58-61; the
    voxel files themselves are not distributed).  Seeded: label 0 outside an elliptic body cylinder, the
    body filled with labels[0], then `n_organs` random ellipsoids (clipped to the body) with labels drawn
    from `labels` (default 1..141, the organ numbers of the ICRP tables)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    labels = np.arange(1, 142) if labels is None else np.asarray(labels)
    out = np.zeros((nz, nrow, ncol), np.uint8)
    yy, xx = np.meshgrid(np.arange(nrow) + 0.5, np.arange(ncol) + 0.5, indexing="ij")
    ax, ay = 0.46 * ncol, 0.46 * nrow
    body = ((xx - ncol / 2) / ax) ** 2 + ((yy - nrow / 2) / ay) ** 2 <= 1.0
    out[:, body] = labels[0]
    for _ in range(n_organs):
        c = rng.uniform([0.2 * ncol, 0.2 * nrow, 0.0], [0.8 * ncol, 0.8 * nrow, nz])
        r = rng.uniform([0.03 * ncol, 0.03 * nrow, 2.0], [0.15 * ncol, 0.2 * nrow, 0.4 * nz])
        lab = labels[rng.integers(0, len(labels))]
        z0, z1 = max(0, int(c[2] - r[2])), min(nz, int(c[2] + r[2]) + 1)
        for z in range(z0, z1):
            w = 1.0 - ((z + 0.5 - c[2]) / r[2]) ** 2
            if w <= 0:
                continue
            m = (((xx - c[0]) / r[0]) ** 2 + ((yy - c[1]) / r[1]) ** 2 <= w) & body
            out[z][m] = lab
    return out


def write_material_tables(dirpath, seed=0, n_media=53, n_organs=140):
    """Seeded media / organ tables in the ICRP CSV formats the converter reads (media rows
    "ID,H,C,N,O,Na,Mg,P,S,Cl,K,Ca,Fe,I" weight %, organ rows "organ,media,density"); medium n_media is air
    (N 70 / O 30, never redrawn).  Returns (media_csv, organs_csv)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    os.makedirs(dirpath, exist_ok=True)
    media, organs = os.path.join(dirpath, "media.csv"), os.path.join(dirpath, "organs.csv")
    with open(media, "w") as f:
        for i in range(1, n_media):
            w = rng.dirichlet(np.ones(13) * 0.5) * 100.0
            f.write(",".join([str(i)] + [f"{v:.1f}" for v in w]) + "\n")
        f.write(f"{n_media},0,0,70,30,0,0,0,0,0,0,0,0,0\n")
    with open(organs, "w") as f:
        for o in range(1, n_organs + 1):
            med = n_media if o == n_organs else int(rng.integers(1, n_media))
            dens = 0.001 if med == n_media else float(rng.uniform(0.9, 2.0))
            f.write(f"{o},{med},{dens:.2f}\n")
    return media, organs


def write_slices(vol, prefix):
    """Writes <prefix>_%03d.raw, one [mat][y][x] uint16 slice per file (P/projectionData.cpp:197)."""
    os.makedirs(os.path.dirname(prefix) or ".", exist_ok=True)
    for z in range(vol.shape[0]):
        np.ascontiguousarray(vol[z], np.uint16).tofile(f"{prefix}_{z:03d}.raw")


# --------------------------------------------------------------------------------------------- spectrum
def spectrum_rows(kvp=100, e0=10, peak_amp=0.3, sigma_kev=1.0, peaks=(59.0, 67.0)):
    """1 keV bins [E, E+1) keV for E = e0 .. kvp (91 bins at 10-100 keV): fraction max(0, (kvp - Ec)/Ec) + A [G(59) + G(67)], normalised."""
    e = np.arange(e0, kvp + 1, dtype=np.float64)
    ec = e + 0.5
    f = np.maximum(0.0, (kvp - ec) / ec)
    for p in peaks:
        if p < kvp:
            f = f + peak_amp * np.exp(-0.5 * ((ec - p) / sigma_kev) ** 2)
    f = f / f.sum()
    return e, f


def write_spectrum_csv(path, kvp=100, e0=10, **kw):
    """The reference's spectrum CSV (P/Spectrum/*.csv): header, then 'lo_MeV,hi_MeV,fraction' rows."""
    e, f = spectrum_rows(kvp, e0, **kw)
    with open(path, "w") as fh:
        fh.write("Energy,1000,Fraction\n")
        for lo, fr in zip(e, f):
            fh.write(f"{lo / 1000:.3f},{(lo + 1) / 1000:.3f},{fr:.6E}\n")
    return len(e)


def klein_nishina_per_electron(e_kev):
    k = np.asarray(e_kev, np.float64) / 510.99895
    l1 = np.log1p(2 * k)
    return 2 * np.pi * R_E_CM ** 2 * ((1 + k) / k ** 2 * (2 * (1 + k) / (1 + 2 * k) - l1 / k) + l1 / (2 * k) -
                                      (1 + 3 * k) / (1 + 2 * k) ** 2)


def write_xraydb(dirpath, seed=1234):
    """photo/comp/rayl_5keVto150keV.csv in the xraydb export format (header H,C,N,O,P,Ca,energy; 29001 rows on a
    5 eV grid; energy in eV).  photo = a_Z Z^4 / E_keV^3 with a_Z = 1 + 0.1 U(-1, 1) (seeded), comp = (Z/A) N_A
    sigma_KN(E), rayl = 0; all cm^2/g."""
    os.makedirs(dirpath, exist_ok=True)
    rng = np.random.Generator(np.random.PCG64(seed))
    a_z = 1.0 + 0.1 * rng.uniform(-1, 1, 6)
    e_ev = 5000.0 + 5.0 * np.arange(29001)
    ek = e_ev / 1000.0
    photo = a_z[None, :] * Z_OF[None, :] ** 4 / ek[:, None] ** 3
    comp = (Z_OF / A_OF)[None, :] * N_A * klein_nishina_per_electron(ek)[:, None]
    rayl = np.zeros_like(photo)
    for name, tab in (("photo", photo), ("comp", comp), ("rayl", rayl)):
        with open(os.path.join(dirpath, f"{name}_5keVto150keV.csv"), "w") as fh:
            fh.write("H,C,N,O,P,Ca,energy\n")
            for r in range(e_ev.size):
                fh.write(",".join(repr(float(v)) for v in tab[r]) + f",{int(e_ev[r])}\n")
    return a_z


def synthetic_spectrum(dirpath, kvp=100, seed=1234):
    """Spectrum and mu/rho tables of the synthetic workloads: writes the xraydb-format CSVs and a kVp spectrum
    CSV (1 keV bins from 10 keV) under dirpath and reads them back through the reference's matching rule
    (attenuation_coefficient).  Returns (Attenuation [6][NE] float32, pdf [NE] float32)."""
    from . import io
    db = os.path.join(dirpath, "db")
    if not os.path.exists(os.path.join(db, "rayl_5keVto150keV.csv")):
        write_xraydb(db, seed)
    spec = os.path.join(dirpath, f"spec{kvp}.csv")
    write_spectrum_csv(spec, kvp=kvp)
    return io.attenuation_coefficient(spec, db)


# --------------------------------------------------------------------------------------------- flexmaps
def flexmap_rows(n_angles, xshift=0.0, yshift=0.0):
    """Angle rows of the shipped flexmaps: 36 -> P/36_little_slide.flexmap (-170.0001 .. 179.99999),
    360 -> P/360_little_slide.flexmap (-180.0001 .. 179.0001)."""
    rows = []
    if n_angles == 36:
        for k in range(36):
            a = -170 + 10 * k
            s = f"{a - 0.0001:.4f}" if a < -90 else f"{a - 0.00001:.5f}"
            rows.append(s)
    elif n_angles == 360:
        for k in range(360):
            a = -180 + k
            rows.append(f"{a - 0.0001:.4f}" if a < 0 else f"{a + 0.0001:.4f}")
    else:
        step = 360.0 / n_angles
        rows = [f"{-180.0 + step * k + 1e-4:.6f}" for k in range(n_angles)]
    return [f"{r} {xshift:g} {yshift:g} 0" for r in rows]


def write_flexmap(path, n_angles, xshift=0.0, yshift=0.0):
    with open(path, "w") as fh:
        fh.write("\n".join(flexmap_rows(n_angles, xshift, yshift)) + "\n")


def flexmap_angles(n_angles, xshift=0.0, yshift=0.0):
    """(angles_rad, xshift, yshift) exactly as the projector reads them (P/projectionData.cpp:164-171)."""
    a = np.array([float(r.split()[0]) for r in flexmap_rows(n_angles)], np.float64)
    return a * 3.141592653589793238462643383279 / 180.0, np.full(a.size, float(xshift)), np.full(a.size, float(yshift))


# --------------------------------------------------------------------------------------------- txt files
def write_proj_txt(path, spectrum, n_ang, W, H, flexmap, prefix, zic, outputname, scale, N, Nz, nmat=6,
                   mat_type="human", xfactor=1000, reg=(0.0, 0.0, 0.0)):
    """Projector info file in the layout of P/txtfile/IR_waterphantom_36pro.txt (6 header lines, 15 blocks)."""
    blocks = [("X-ray Energy Spectrum", spectrum), ("total projections scanned", n_ang), ("projection width", W),
              ("projection height", H), ("angle data file", flexmap), ("Material density for input", prefix),
              ("IC z position(slice number)", zic), ("output projection name", outputname),
              ("scale for x, y, z", f"{scale[0]:f}  {scale[1]:f}  {scale[2]:f}"),
              ("input size (square format)", N), ("num of slices", Nz), ("mat_num", nmat),
              ("material type (human, gammex, catphan, etc)", mat_type),
              ("Factor for density value", xfactor),
              ("prior image shift (-x, -y, -z) (opposite of Couch Shift direction)", f"{reg[0]} {reg[1]} {reg[2]}")]
    with open(path, "w") as fh:
        fh.write("////////\n//\n//    info file for the virtual CBCT projector\n//    keep the header, blank and "
                 "comment lines\n//\n////////\n")
        for c, v in blocks:
            fh.write(f"\n// {c}\n{v}\n")


def write_fdk_txt(path, n_ang, W, H, proj_path, ext, flexmap, R, zstart_cm, step, Z, scale, xfactor, output):
    """FDK info file in the layout of F/txtfile/FDK_36pro_water20cm.txt (6 header lines, 18 blocks)."""
    blocks = [("X-ray Energy Spectrum", "kV"), ("total projections scanned", n_ang), ("projection width", W),
              ("projection height", H), ("projection image file header bytes", 0),
              ("projection image data directory for input", proj_path), ("projection image file number digits", 8),
              ("first projection image file number in HEX", "hore"), ("projection image file extension", ext),
              ("angle data file", flexmap), ("reconstruction size (square format)", R),
              ("reconstruction target class string", "*"), ("reconstruction start z [cm]", int(zstart_cm)),
              ("reconstruction z step[cm]", step), ("reconstruction slices", Z),
              ("reconstruction size factor [cm]", f"{scale:f}"), ("Factor for output CT value", int(xfactor)),
              ("reconstructed data file for output", output)]
    with open(path, "w") as fh:
        fh.write("////////\n//\n//    info file for CBCT reconstruction\n//    keep the header, blank and comment "
                 "lines\n//\n////////\n")
        for c, v in blocks:
            fh.write(f"\n// {c}\n{v}\n")


# --------------------------------------------------------------------------------------------- workloads
WORKLOADS = {
    # water cylinder 20 cm, 512^2 x 64 at 0.5 mm / 1 mm, 36 angles, 400^2 detector, 91 bins; FDK 270^2 x 64
    "c1": dict(N=512, Nz=64, s=0.05, t=0.1, radius=10.0, n_ang=36, H=400, W=400, xshift_proj=0.0,
               xshift_fdk=0.0, R=270, Z=64, zstart=-3, zstep=0.1, rscale=0.1),
    # water cylinder 25 cm, 512^2 x 128, 360 angles, 115 mm half-fan offset; FDK 270^2 x 128
    "c2": dict(N=512, Nz=128, s=0.05, t=0.1, radius=12.5, n_ang=360, H=400, W=400, xshift_proj=1.15,
               xshift_fdk=11.5, R=270, Z=128, zstart=-6, zstep=0.1, rscale=0.1),
}
