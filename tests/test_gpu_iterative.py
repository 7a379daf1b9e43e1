"""GPU parity of the iterative-reconstruction kernels (P/lengthproj.cpp, P/IR_ImageCBCT.cpp) through the C ABI:
bit-exact against the FP64 oracle and against the reference's own outputs (fixture)."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from virtualcbct_amd import iterative
from virtualcbct_amd._lib import VcbctError

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "iterative_small.npz")


@pytest.fixture(scope="module")
def torch():
    import torch as T
    if not T.cuda.is_available():
        pytest.skip("no GPU")
    return T


def make_projector(g):
    from virtualcbct_amd.projector import Projector
    return Projector(g.N, g.s, g.Nz, g.t, g.H, g.W, g.n_ang, g.cos_t, g.sin_t, g.xc, g.yc, g.x_reg, g.y_reg, g.z_reg,
                     g.det_pitch, g.det_pitch_iso, g.sad, g.det_pixel_num, g.nmat)


def geom(N, Nz, s, t, H, W, angles_deg, x_reg=0.0, y_reg=0.0, z_reg=0.0):
    ang = np.deg2rad(np.asarray(angles_deg, float))
    cs, sn, xc, yc = O.trig_table_proj(W, H, ang, np.full(len(ang), 0.013), np.full(len(ang), -0.02))
    return O.Geometry(N=N, s=s, Nz=Nz, t=t, H=H, W=W, cos_t=cs, sin_t=sn, xc=xc, yc=yc, x_reg=x_reg, y_reg=y_reg,
                      z_reg=z_reg)


CASES = [
    (24, 6, 0.5, 0.4, 20, 28, [-170.0001, -90.00001, 0.0, 33.3, 90.0, 179.99999], 0.0, 0.0, 0.0),
    (17, 5, 0.3, 0.7, 31, 25, [-45.0, 10.0, 120.5, 200.0], 0.3, 0.1, -0.2),
    (32, 4, 1.1, 0.2, 16, 16, list(np.arange(-180, 180, 20.0) + 0.5), 0.0, 0.0, 0.9),
    (96, 20, 0.25, 0.3, 64, 80, list(np.linspace(-180, 170, 36)), 0.05, -0.1, 0.4),
]


@pytest.mark.parametrize("case", CASES)
def test_lengthproj_matches_oracle(torch, case):
    N, Nz, s, t, H, W, angs, xr, yr, zr = case
    g = geom(N, Nz, s, t, H, W, angs, xr, yr, zr)
    rng = np.random.default_rng(N * 7 + Nz)
    proj = rng.integers(0, 65536, (len(angs), H, W)).astype(np.uint16)
    P = make_projector(g)
    dproj = torch.from_numpy(proj.view(np.int16)).cuda()
    for ite, a0, a1 in ((0, 0, None), (1, 0, None), (0, 1, 3)):
        got = iterative.lengthproj(P, dproj, ite, a0, a1).cpu().numpy()
        ref = O.lengthproj(g, proj, ite, a0, g.n_ang if a1 is None else a1)
        assert np.array_equal(got, ref), (case[:6], ite, a0, a1, np.abs(got - ref).max())
    P.close()


def test_lengthproj_and_update_match_reference_fixture(torch):
    d = np.load(GOLD)
    g = O.Geometry(N=int(d["N"]), s=float(d["s"]), Nz=int(d["Nz"]), t=float(d["t"]), H=int(d["H"]), W=int(d["W"]),
                   cos_t=d["cos"], sin_t=d["sin"], xc=d["xc"], yc=d["yc"], x_reg=float(d["x_reg"]),
                   z_reg=float(d["z_reg"]))
    P = make_projector(g)
    dproj = torch.from_numpy(d["proj"].view(np.int16)).cuda()
    assert np.array_equal(iterative.lengthproj(P, dproj, 0).cpu().numpy(), d["npf"])
    assert np.array_equal(iterative.lengthproj(P, dproj, 1).cpu().numpy(), d["npf_ite1"])
    out = iterative.IR_ImageCBCT(torch.from_numpy(d["before"]).cuda(), torch.from_numpy(d["npf"]).cuda(),
                                 torch.from_numpy(d["npf_re"]).cuda(), float(d["s"]), 0.3, 0.05)
    assert np.array_equal(out.cpu().numpy(), d["ir"])
    P.close()


@pytest.mark.parametrize("N,Nz,s", [(24, 6, 0.5), (13, 4, 1.3), (130, 9, 0.1)])
def test_ir_update_matches_oracle(torch, N, Nz, s):
    rng = np.random.default_rng(N)
    before = rng.integers(-50, 3000, (Nz, N, N)).astype(np.int16)
    before[0, 1, 1] = 32000  # a clamp to 5000 and, at its neighbour, an int-overflowed TV term (NaN)
    npf = rng.uniform(1e3, 1e6, (Nz, N, N)).astype(np.float32)
    npr = rng.uniform(1e3, 1e6, (Nz, N, N)).astype(np.float32)
    npr[0, 1, 1] = 1e7
    db, dn, dr = (torch.from_numpy(a).cuda() for a in (before, npf, npr))
    for w, tv, pr in ((0.3, 0.05, 0.0), (0.0, 0.0, 0.0), (0.9, 2.0, 1.0)):
        got = iterative.IR_ImageCBCT(db, dn, dr, s, w, tv, pr).cpu().numpy()
        assert np.array_equal(got, O.ir_update(before, npf, npr, s, w, tv, pr), equal_nan=True), (N, w, tv)


def test_lengthproj_argument_checks(torch):
    g = geom(16, 3, 0.5, 0.5, 12, 12, [10.0, 20.0])
    P = make_projector(g)
    proj = torch.zeros((2, 12, 12), dtype=torch.int16, device="cuda")
    with pytest.raises(VcbctError, match="angle range"):
        iterative.lengthproj(P, proj, 0, 1, 3)
    with pytest.raises(ValueError):
        iterative.lengthproj(P, torch.zeros((1, 12, 12), dtype=torch.int16, device="cuda"))
    P.close()
