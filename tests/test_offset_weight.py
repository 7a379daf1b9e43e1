"""Offset-detector (half-fan) weighting: the oracle's restatement pinned by the properties that define it
(no reference counterpart: the reference computes the size class and angle mask for it, F/main.cpp:57, and
never applies a weighting).  The GPU kernel is compared with this oracle in test_gpu_fdk.py."""
import numpy as np
import pytest

from oracle import oracle as O

W = 400
# panel shifts (cm at the detector; 0.1 cm per pixel at W = 400): the c2 half-fan offset, a centred panel,
# a shift to the other side, and a half-pixel shift
SHIFTS = [11.5, 0.0, -3.0, 2.05]


def _centre(xshift):
    return W / 2 + xshift / (0.05 * 800 / W) - 0.5


@pytest.mark.parametrize("xs", SHIFTS)
@pytest.mark.parametrize("taper", [0.0, 40.0])
def test_conjugate_columns_sum_to_one(xs, taper):
    """w(t) + w(-t) = 1 for the columns placed symmetrically about the central ray (the redundancy
    condition: each ray measured twice in a 360-degree scan is counted once in total)."""
    e = O.offset_exponents(W, xs, taper)[0]
    c = _centre(xs)
    twice = 2 * c
    assert abs(twice - round(twice)) < 1e-9  # the chosen shifts put the central ray on a column or edge
    lo, hi = int(np.floor(c)), int(np.ceil(c))
    if lo == hi:  # central ray on a column centre: w = 1/2 there
        assert e[lo] == pytest.approx(1.0, abs=1e-15)
    pairs = 0
    for d in range(0, W):
        a, b = lo - d, hi + d
        if a < 0 or b >= W:
            break
        if a == b:
            continue
        assert e[a] + e[b] == pytest.approx(2.0, abs=1e-12), (a, b)
        pairs += 1
    assert pairs > 10


@pytest.mark.parametrize("xs", SHIFTS)
def test_long_side_full_weight_short_side_tapered(xs):
    e = O.offset_exponents(W, xs)[0]
    sh = xs / 0.1
    delta = W / 2 - abs(sh)
    c = _centre(xs)
    t = (-1.0 if sh > 0 else 1.0) * (np.arange(W) - c)
    assert np.all(e[t >= delta] == 2.0)
    inside = np.abs(t) < delta
    assert np.all((e[inside] > 0) & (e[inside] < 2))
    # non-decreasing towards the long side
    order = np.argsort(t)
    assert np.all(np.diff(e[order]) >= -1e-15)
    # every column lies within the overlap's outer edge (the short edge is delta away from the central ray)
    assert np.all(t > -delta)


def test_closed_form_values():
    # c = 314.5 for the c2 offset, long side towards column 0, tau = 85
    e = O.offset_exponents(W, 11.5)[0]
    c = 314.5
    for col in (250, 290, 314, 315, 350, 399):
        t = -(col - c)
        assert e[col] == pytest.approx(2 * np.sin(np.pi / 4 * (1 + t / 85.0)) ** 2, rel=1e-14, abs=1e-15)
    # half-way into the long side: 2 sin^2(3 pi / 8) = 1 + sqrt(2)/2; a narrow taper
    e2 = O.offset_exponents(W, 0.0, taper=20.0)[0]  # c = 199.5, long side towards higher columns
    assert e2[209] == pytest.approx(2 * np.sin(np.pi / 4 * (1 + 9.5 / 20)) ** 2, rel=1e-14)
    assert e2[179] == 0.0 and e2[220] == 2.0 and 0.0 < e2[180] and e2[219] < 2.0


def test_weight_acts_on_the_line_integral_and_pads_the_short_side():
    rng = np.random.default_rng(5)
    q = rng.uniform(0.0, 4.0, (3, 5, W))
    p = np.exp(-q).astype(np.float32)
    for xs in (np.array([11.5, 11.5, 11.43]), np.array([-3.0, -2.9, -3.0]), np.zeros(3)):
        ys = np.array([0.0, 0.3, -1.0])
        out, xo, yo = O.offset_weight(p, xs, ys)
        sh = xs / 0.1
        P = int(round(2 * np.abs(sh).max()))
        assert out.shape == (3, 5, W + P)
        off = P if (xs < 0).any() else 0
        real = out[:, :, off:off + W]
        e = O.offset_exponents(W, xs)
        q_in = -np.log(p.astype(np.float64))
        np.testing.assert_allclose(-np.log(real.astype(np.float64)), e[:, None, :] * q_in, rtol=2e-6, atol=1e-6)
        full = e[:, None, :].repeat(5, 1) == 2.0
        np.testing.assert_array_equal(real[full], (p.astype(np.float64) ** 2).astype(np.float32)[full])
        pad = np.ones(W + P, bool)
        pad[off:off + W] = False
        assert np.all(out[:, :, pad] == 1.0)  # no attenuation in the padding
        # the central ray keeps its column on the padded panel: Wo/2 + shift' - 0.5 = c + off (pixels of the
        # same size: shift' is given in the padded panel's units, 0.05 * 800 / Wo)
        Wo = W + P
        np.testing.assert_allclose(Wo / 2 + xo / (0.05 * 800 / Wo) - 0.5, W / 2 + sh - 0.5 + off, atol=1e-9)
        # and the filter's row centre, which it takes from the panel width: -Wo/2 - yshift' = -W/2 - yshift
        np.testing.assert_allclose(-Wo / 2 - yo / (0.05 * 800 / Wo), -W / 2 - ys / 0.1, atol=1e-9)
        if (xs > 0).any():
            # symmetric about the central ray: as many columns on each side of it
            c = W / 2 + sh.max() - 0.5
            assert abs((c + 0.5) - (Wo - c - 0.5)) <= 1.0


def test_shifts_of_both_signs_or_without_overlap_are_rejected():
    with pytest.raises(ValueError):
        O.offset_weight(np.ones((1, 2, W), np.float32), [20.0], [0.0])
    with pytest.raises(ValueError):
        O.offset_weight(np.ones((2, 2, W), np.float32), [3.0, -3.0], [0.0, 0.0])
    assert O.offset_exponents(W, 25.0)[0][0] == -1.0
    assert O.half_fan_width(W, [11.5, 11.0]) == W + 230


def test_half_fan_reconstruction_matches_centred_panel():
    """The oracle's FDK of a half-fan scan (c2's 115-pixel offset) with the weighting and padding matches the
    centred panel's (monochromatic spectrum, so a flat 25 cm water cylinder); with the weighting but without
    the padding, the region that only the long side sees comes out far too high (the filtered projection's
    tails beyond the short edge are lost).  Reduced sizes: 128^2 x 8 voxels of 2 mm, 60 angles; the rows
    around the central plane only."""
    import tempfile
    from virtualcbct_amd import io, phantoms
    N, s, Nz, t, H, n_ang, R, rs = 128, 0.2, 8, 0.2, W, 60, 135, 0.2
    with tempfile.TemporaryDirectory() as td:
        phantoms.write_xraydb(f"{td}/db")
        phantoms.write_spectrum_csv(f"{td}/s.csv")
        atten, pdf = io.attenuation_coefficient(f"{td}/s.csv", f"{td}/db")
        mono = np.zeros_like(pdf)
        mono[50] = 1.0
        vol = phantoms.water_cylinder(N, Nz, s, 12.5)
        yy, xx = np.mgrid[0:R, 0:R]
        r = np.hypot(xx - (R - 1) / 2, yy - (R - 1) / 2) * rs
        disk, outer = r < 11.0, (r > 6.5) & (r < 11.0)
        recs = {}
        for name, xp, xf in (("centred", 0.0, 0.0), ("half", 1.15, 11.5)):
            fm = f"{td}/{name}.flexmap"
            phantoms.write_flexmap(fm, n_ang, xf)
            a_f, xs_f, ys_f = io.read_fdk_angles(fm, n_ang)
            ang, xs, ys = phantoms.flexmap_angles(n_ang, xp)
            cs, sn, xc, yc = O.trig_table_proj(W, H, ang, xs, ys)
            g = O.Geometry(N, s, Nz, t, H, W, cs, sn, xc, yc)
            p = np.ascontiguousarray(O.forward_collapsed(g, vol, atten, mono, 1000.0)[:, 190:210])
            if name == "centred":
                f = O.filter_sinogram(p, xs_f, ys_f)
                recs[name] = O.reconstruct_image(R, 1, 0, 0.1, rs, f, 1000, a_f, xs_f)[0]
                continue
            pw, xo, yo = O.offset_weight(p, xs_f, ys_f)
            pit = 0.032509752 * pw.shape[2] / W
            f = O.filter_sinogram(pw, xo, yo, pitch_iso=pit)
            recs["padded"] = O.reconstruct_image(R, 1, 0, 0.1, rs, f, 1000, a_f, xo, pitch_iso=pit)[0]
            crop = np.ascontiguousarray(pw[:, :, :W])  # weighted, not padded
            f = O.filter_sinogram(crop, xs_f, ys_f)
            recs["unpadded"] = O.reconstruct_image(R, 1, 0, 0.1, rs, f, 1000, a_f, xs_f)[0]
    # ring means (1 cm rings out to 11 cm): the level the weighting must keep; the pixel-wise streaks of
    # 60 views differ between the two panel positions and are not compared
    ref = recs["centred"]
    level = float(ref[disk].mean())

    def rings(img):
        return np.array([img[(r >= a) & (r < a + 1)].mean() for a in range(11)])

    dev = {k: float(np.abs(rings(recs[k]) - rings(ref)).max()) / level for k in ("padded", "unpadded")}
    out_u = float(recs["unpadded"][outer].mean()) / float(ref[outer].mean())
    print(f"level {level:.1f}; largest ring-mean difference / level: padded {dev['padded']:.4f}, "
          f"unpadded {dev['unpadded']:.4f}; unpadded outer ring / centred {out_u:.3f}")
    assert dev["padded"] < 0.02
    assert out_u > 1.1
