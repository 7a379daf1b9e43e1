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


def _oracle_mlem(g, mu, b, x0, iters, frac, wbefore, weight_tv):
    """The same loop with the FP64 oracle: A from orc_logproj (one bin), A^T from orc_adjoint (the gradient of
    sum (p - t)^2 at t = p - r / 2 is A^T r), the update from orc_ir_update; int16 truncation between."""
    pdf = np.ones(1, np.float32)
    at = np.asarray(mu, np.float32).reshape(-1, 1)
    f = np.asarray(frac, np.float32).reshape(1, -1, 1, 1)

    def maps_of(x):  # [z][mat][y][x]
        return np.ascontiguousarray(f * (x.astype(np.float32)[:, None] * np.float32(1e-3)))

    def AT(x, r):
        m = maps_of(x)
        p = O.logproj(g, m, at, pdf, 1.0)
        gr, _ = O.adjoint(g, m, at, pdf, 1.0, p - r / 2.0)
        return (gr * f).sum(1) * 1e-3

    npf_re = AT(np.zeros_like(x0), b).astype(np.float32)
    x = x0.astype(np.int16)
    out = None
    for _ in range(iters):
        ax = O.logproj(g, maps_of(x), at, pdf, 1.0)
        npf = AT(x, ax).astype(np.float32)
        out = O.ir_update(x, npf, npf_re, g.s, wbefore, weight_tv)
        x = np.trunc(out).astype(np.int16)
    return out, x


def test_mlem_tv_loop_on_the_kernels_matches_oracle_loop(torch):
    """The MLEM + TV loop built on the forward and adjoint kernels (iterative.MlemTV) against the same loop run
    with the FP64 oracle: equal images to the int16 truncation after three iterations, a single scatter channel,
    and the data misfit falling."""
    from virtualcbct_amd.iterative import MlemTV, WATER_FRACTIONS
    N, Nz = 24, 6
    g = geom(N, Nz, 0.5, 0.5, 20, 20, list(np.arange(-180, 180, 30.0) + 0.37))
    yy, xx = np.mgrid[0:N, 0:N]
    r = np.hypot(xx - (N - 1) / 2, yy - (N - 1) / 2)
    truth = np.zeros((Nz, N, N), np.int16)
    truth[:, r < 10] = 1000
    truth[2:4, (r < 4)] = 1800
    mu = np.array([0.33, 0.2, 0.19, 0.18, 0.17, 0.2])
    at = mu.astype(np.float32).reshape(-1, 1)
    f = np.asarray(WATER_FRACTIONS, np.float32).reshape(1, -1, 1, 1)
    b = O.logproj(g, np.ascontiguousarray(f * (truth.astype(np.float32)[:, None] * np.float32(1e-3))), at,
                  np.ones(1, np.float32), 1.0).astype(np.float32)
    x0 = np.where(r < 10, 800, 0).astype(np.int16)[None].repeat(Nz, 0)
    P = make_projector(g)
    solver = MlemTV(P, mu, torch.from_numpy(b).cuda(), wbefore=0.0, weight_tv=1e-5)
    assert P.adjoint_channels == 1
    out, xg = solver.run(torch.from_numpy(x0).cuda(), 3)
    oo, xo = _oracle_mlem(g, mu, b, x0, 3, WATER_FRACTIONS, 0.0, 1e-5)
    d = np.abs(xg.cpu().numpy().astype(np.int32) - xo.astype(np.int32))
    rel = float(np.abs(out.cpu().numpy() - oo).max() / np.abs(oo).max())
    print(f"MLEM-TV, 3 iterations: int16 images differ by at most {d.max()} ({int((d > 0).sum())} voxels), "
          f"FP32 images {rel:.2e} of max")
    assert d.max() <= 1 and (d > 0).mean() < 0.01
    assert rel < 1e-4
    # the misfit falls
    x = torch.from_numpy(x0).cuda()
    res = []
    for _ in range(4):
        res.append(float((solver.A(x) - solver.npf_re.new_tensor(b)).norm()))
        x = solver.step(x).to(torch.int16)
    assert res[-1] < res[0]
    P.close()
