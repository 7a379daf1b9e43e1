"""Organ-label phantom -> elemental maps (ICRPphantoms_2mmslice/To_material_distribution_for512_random.cpp):
the oracle against the reference program's own outputs, and the product's host table builder
(vcbct_material_table, no device work) against the oracle."""
import ctypes as C
import hashlib
import os

import numpy as np
import pytest

from oracle import oracle as O
from virtualcbct_amd import material_maps as MM
from virtualcbct_amd import phantoms
from virtualcbct_amd._lib import VcbctError

HERE = os.path.dirname(os.path.abspath(__file__))
ICRP = os.path.join(HERE, "golden", "icrp")
FIX = os.path.join(HERE, "golden", "material_maps.npz")
TABLES = {s: (os.path.join(ICRP, f"{s}_media.csv"), os.path.join(ICRP, f"{s}_organs.csv")) for s in ("AM", "AF")}


def fixture_labels(seed):
    # the recipe of tests/golden/gen_golden.py (material_maps_case)
    lab = phantoms.synthetic_labels(543, 271, 100, seed=seed)
    lab[0, 0, :] = 0
    lab[0, 0, :256] = np.arange(256, dtype=np.uint8)
    return lab


def oracle_lut(sex, randomize=True, seed=1):
    return O.material_lut(*TABLES[sex], randomize=randomize, seed=seed)


# ------------------------------------------------------------------------------------ the oracle itself

def test_rand_restatement_matches_libc():
    libc = C.CDLL("libc.so.6")
    for seed in (0, 1, 2, 77, 2 ** 31 - 1, 4294967295):
        libc.srand(C.c_uint(seed))
        seq = np.array([libc.rand() for _ in range(1500)], np.int32)
        assert np.array_equal(seq, O.rand_sequence(seed, 1500)), seed


def test_oracle_matches_reference_fixture():
    """The reference program's outputs (fixture, from oracle/_ref/to_material on the seeded labels): its
    printed media table, its label -> map table and digests of every per-slice file."""
    fx = np.load(FIX)
    media, organs = TABLES["AM"]
    w = O.material_weights(media, True, 1)
    ids = np.array([int(r.split(",")[0]) for r in open(media).read().split()])
    printed = np.array([[float(f"{v:f}") for v in row] for row in w])
    assert np.array_equal(fx["media_printed"][:, 0], ids)
    assert np.array_equal(fx["media_printed"][:, 1:], printed)
    assert np.array_equal(oracle_lut("AM"), fx["lut"])
    lab = fixture_labels(int(fx["seed"]))
    md = O.labels_to_maps(lab, w, O.organ_table(organs))
    assert np.array_equal(md.astype(np.int64).sum(axis=(2, 3)), fx["md_sums"])
    assert hashlib.sha256(md.tobytes()).hexdigest() == fx["md_sha256"].item().decode()
    whole = np.ascontiguousarray(md.transpose(1, 0, 2, 3))
    assert hashlib.sha256(whole.tobytes()).hexdigest() == fx["whole_sha256"].item().decode()


@pytest.mark.skipif(not O.ref_to_material_available(), reason="reference converter not built (oracle/_ref)")
def test_oracle_matches_reference_program(tmp_path):
    """A second label volume, run through the reference program itself: test.raw and all 100 slice files
    bit-exact, and its printed weight table."""
    rng = np.random.default_rng(9)
    lab = fixture_labels(777)
    lab[50:53] = rng.integers(0, 256, lab[50:53].shape).astype(np.uint8)  # every label, unmatched ones too
    media, organs = TABLES["AM"]
    whole, md, stdout = O.ref_to_material_run(lab, media, organs, str(tmp_path))
    w = O.material_weights(media, True, 1)
    mine = O.labels_to_maps(lab, w, O.organ_table(organs))
    assert np.array_equal(mine, md)
    assert np.array_equal(mine.transpose(1, 0, 2, 3), whole)
    lines = stdout.splitlines()
    i0 = lines.index("AM_media.csv ") + 1
    for row, ln in zip(w, lines[i0:i0 + len(w)]):
        assert " ".join(f"{v:f}" for v in row) == " ".join(ln.split()[1:])


def test_oracle_unrandomized_closed_form():
    """Without the draws the weights are the CSV's H C N P Ca, O = 100 minus those five (FP32), and a
    label's value is trunc(w x density x 10) of its last organ row."""
    for sex in ("AM", "AF"):
        media, organs = TABLES[sex]
        rows = np.array([[float(v) for v in r.split(",")] for r in open(media).read().split()], np.float32)
        w = O.material_weights(media, False)
        five = rows[:, [1, 2, 3, 7, 11]]
        assert np.array_equal(w[:, [0, 1, 2, 4, 5]], five)
        o = np.float32(100) - five[:, 0] - five[:, 1] - five[:, 2] - five[:, 3] - five[:, 4]
        assert np.array_equal(w[:, 3], np.maximum(o, 0)) and (o >= 0).all()
        lut = oracle_lut(sex, randomize=False)
        org = np.array([[float(v) for v in r.split(",")] for r in open(organs).read().split()])
        expect = np.zeros((256, 6), np.int16)
        for lab_, med, dens in org:
            expect[int(lab_)] = np.trunc(w[int(med) - 1] * np.float32(dens) * np.float32(10)).astype(np.int16)
        assert np.array_equal(lut, expect)


def test_oracle_draw_statistics():
    """The redrawn H, C, N weights over many seeds: mean w and standard deviation 0.1 w (Box-Muller)."""
    media, _ = TABLES["AM"]
    w0 = O.material_weights(media, False)
    ws = np.stack([O.material_weights(media, True, s) for s in range(1, 801)])
    for k in (0, 1, 2):
        sel = w0[:-1, k] > 1.0  # away from the clamp at 0
        m = ws[:, :-1, k][:, sel].mean(0)
        sd = ws[:, :-1, k][:, sel].std(0)
        assert np.allclose(m, w0[:-1, k][sel], rtol=0.02)
        assert np.allclose(sd, 0.1 * w0[:-1, k][sel], rtol=0.12)
    assert np.array_equal(ws[:, -1], np.broadcast_to(w0[-1], ws[:, -1].shape))  # medium 53 is never redrawn


# ------------------------------------------------------------------------- the product's table builder

@pytest.mark.parametrize("sex", ["AM", "AF"])
@pytest.mark.parametrize("randomize,seed", [(True, 1), (True, 2), (True, 12345), (False, 1)])
def test_material_table_matches_oracle(sex, randomize, seed):
    lut, w = MM.material_table(*TABLES[sex], randomize=randomize, seed=seed)
    assert lut.shape == (256, 8) and not lut[:, 6:].any()
    assert np.array_equal(w, O.material_weights(TABLES[sex][0], randomize, seed))
    assert np.array_equal(lut[:, :6].view(np.int16), oracle_lut(sex, randomize, seed))


def test_material_table_reference_fixture():
    lut, _ = MM.material_table(*TABLES["AM"])
    assert np.array_equal(lut[:, :6].view(np.int16), np.load(FIX)["lut"])


def test_material_table_leaves_process_rand_alone():
    libc = C.CDLL("libc.so.6")
    libc.srand(5)
    a = [libc.rand() for _ in range(3)]
    libc.srand(5)
    libc.rand()
    MM.material_table(*TABLES["AM"], seed=9)
    assert [libc.rand() for _ in range(2)] == a[1:]


def test_material_table_last_row_wins_and_wide_labels(tmp_path):
    media = tmp_path / "m.csv"
    media.write_text("1,10,20,5,60,0,0,1,0,0,0,2,0,0\n2,5,5,5,5,0,0,50,0,0,0,60,0,0\n")
    organs = tmp_path / "o.csv"
    organs.write_text("3,1,1.0\n3,2,2.0\n300,1,1.0\n7,2,0.5\n")
    lut, w = MM.material_table(media, organs, randomize=False)
    # medium 2 sums past 100: O = 0 and the five are scaled to 100
    assert w[1, 3] == 0 and abs(float(w[1].sum()) - 100.0) < 1e-4
    assert np.array_equal(lut[:, :6].view(np.int16), O.material_lut(media, organs, False))
    assert lut[3, 0] == int(np.float32(w[1, 0]) * np.float32(2.0) * np.float32(10))  # the second row for label 3
    assert not lut[np.setdiff1d(np.arange(256), [3, 7])].any()  # label 300 matches no 8-bit label


def test_material_table_errors(tmp_path):
    media, organs = TABLES["AM"]
    with pytest.raises(VcbctError, match="not found"):
        MM.material_table(tmp_path / "none.csv", organs)
    bad = tmp_path / "bad.csv"
    bad.write_text("1,2,3\n")
    with pytest.raises(VcbctError, match="malformed"):
        MM.material_table(bad, organs)
    org = tmp_path / "o.csv"
    org.write_text("1,99,1.0\n")
    with pytest.raises(VcbctError, match="medium 99"):
        MM.material_table(media, org)
    dense = tmp_path / "d.csv"
    dense.write_text("1,1,500.0\n")
    with pytest.raises(VcbctError, match="16-bit"):
        MM.material_table(media, dense)


def test_synthetic_labels_shape_and_labels():
    lab = phantoms.synthetic_labels(60, 30, 8, seed=3, n_organs=10)
    assert lab.shape == (8, 30, 60) and lab.dtype == np.uint8
    assert lab[:, 0, 0].max() == 0 and set(np.unique(lab)) <= set(range(142))
    assert np.array_equal(lab, phantoms.synthetic_labels(60, 30, 8, seed=3, n_organs=10))
