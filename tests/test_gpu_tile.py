"""Tile shapes of the forward launches: every shape (row groups x columns per workgroup, order, XCD split) is a
schedule only, so the intensity and log projections must be bit-identical under all of them, on a geometry
whose row, column and angle counts divide none of the tile sizes; the autotuner's pick is one of the shapes
and is logged; invalid shapes are refused."""
import itertools
import os

import numpy as np
import pytest

from oracle import oracle as O
from virtualcbct_amd import phantoms

pytestmark = pytest.mark.gpu

SHAPES = [(1, 2), (2, 1), (1, 4), (2, 2), (4, 1), (1, 8), (2, 4), (4, 2)]


@pytest.fixture(scope="module")
def torch():
    import torch as T
    if not T.cuda.is_available():
        pytest.skip("no GPU")
    return T


@pytest.fixture(scope="module")
def case():
    # 150 rows = 3 row groups (the last one partial), 70 columns, 13 angles
    N, Nz, H, W, nA = 64, 24, 150, 70, 13
    ang = np.linspace(0, 2 * np.pi, nA, endpoint=False)
    cos_t, sin_t, xc, yc = O.trig_table_proj(W, H, ang, np.zeros(nA), np.zeros(nA))
    g = O.Geometry(N=N, s=0.25, Nz=Nz, t=0.25, H=H, W=W, cos_t=cos_t, sin_t=sin_t, xc=xc, yc=yc, z_reg=0.1)
    vol = phantoms.ellipsoid_phantom(N, Nz, 0.25, 0.25, seed=7)
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "reproj_small.npz"))
    return g, vol, d["atten"], d["pdf"]


def _projector(g):
    from virtualcbct_amd.projector import Projector
    return Projector(g.N, g.s, g.Nz, g.t, g.H, g.W, g.n_ang, g.cos_t, g.sin_t, g.xc, g.yc, g.x_reg, g.y_reg, g.z_reg,
                     g.det_pitch, g.det_pitch_iso, g.sad, g.det_pixel_num, g.nmat, 0)


def test_every_tile_shape_gives_identical_intensity_and_logproj(torch, case):
    g, vol, atten, pdf = case
    P = _projector(g)
    P.set_volume(vol)
    P.set_spectrum(atten, pdf)
    maps = torch.from_numpy(np.ascontiguousarray(vol.transpose(1, 0, 2, 3).astype(np.float32)) * 1e-3).cuda()
    Q = _projector(g)
    Q.set_volume(maps)
    Q.set_spectrum(atten, pdf)
    P.set_tile(1, 4, 0, 0)
    ref_i = P.forward(1000.0).clone()
    ref_part = P.forward(1000.0, a0=4, a1=7).clone()   # 3 angles: most XCDs idle under split 1
    Q.set_tile(1, 4, 0, 0)
    ref_p = Q.logproj(1.0).clone()
    ref_ps, L0 = Q.logproj_save(1.0)
    assert torch.isfinite(ref_i).all() and (ref_i > 0).any() and torch.isfinite(ref_p).all()
    assert (ref_p > 0).any()
    for (rg, c), order, split in itertools.product(SHAPES, (0, 1), (0, 1)):
        print(f"tile {rg}x{c} order {order} split {split}", flush=True)
        P.set_tile(rg, c, order, split)
        a = P.forward(1000.0)
        assert P.tile == (rg, c, order, split)
        assert torch.equal(a, ref_i)
        part = torch.zeros_like(a)
        P.forward(1000.0, out=part, a0=4, a1=7)
        assert torch.equal(part[4:7], ref_part[4:7]) and not part[:4].any() and not part[7:].any()
        Q.set_tile(rg, c, order, split)
        assert torch.equal(Q.logproj(1.0), ref_p)
        ps, L = Q.logproj_save(1.0)
        assert torch.equal(ps, ref_ps) and torch.equal(L, L0)
    P.close()
    Q.close()


def test_autotuned_shape_is_a_candidate_and_logged(torch, case):
    from virtualcbct_amd.batch import autotune_log
    g, vol, atten, pdf = case
    P = _projector(g)
    P.set_volume(vol)
    P.set_spectrum(atten, pdf)
    P.set_tile(1, 4, 0, 0)
    ref = P.forward(1000.0).clone()
    P.set_tile(0)
    a = P.forward(1000.0)
    rg, c, order, split = P.tile
    print("autotuned tile:", P.tile)
    assert (rg, c) in SHAPES and order in (0, 1) and split in (0, 1)
    assert torch.equal(a, ref)
    log = autotune_log()
    print(log)
    assert f"fwd dev0 N{g.N} Nz{g.Nz} H{g.H} W{g.W} nA{g.n_ang}" in log
    assert f"-> {rg}x{c}/o{order}/s{split}" in log
    # the second launch reuses the cached choice
    b = P.forward(1000.0)
    assert P.tile == (rg, c, order, split) and torch.equal(b, ref)
    P.close()


def test_adjoint_under_every_four_wave_shape(torch, case):
    """The adjoint's coefficient kernel and scatter under the 12 shapes of 4 waves: the loss is the same sum of
    the same per-ray FP64 terms (reordered), the gradient the same FP32 additions in another order; the
    backward from saved line integrals likewise, and the autotuned choice agrees too."""
    g, vol, atten, pdf = case
    maps = torch.from_numpy(np.ascontiguousarray(vol.transpose(1, 0, 2, 3).astype(np.float32)) * 1e-3).cuda()
    Q = _projector(g)
    Q.set_volume(maps)
    Q.set_spectrum(atten, pdf)
    target = Q.logproj(1.0).clone() * 0.9
    gp = torch.linspace(-1.0, 1.0, Q.n_rays, device="cuda").reshape(Q.shape_proj)
    _, L = Q.logproj_save(1.0)
    Q.set_adjoint_tile(1, 4, 0, 0)
    g0, l0 = Q.adjoint(target, 1.0)
    b0 = Q.logproj_backward_saved(L, gp, 1.0)
    scale_g, scale_b = float(g0.abs().max()), float(b0.abs().max())
    assert scale_g > 0 and scale_b > 0 and float(l0) > 0
    shapes = [(r, c, o, s) for (r, c) in [(1, 4), (2, 2), (4, 1)] for o in (0, 1) for s in (0, 1)] + [(0, 0, 0, 0)]
    for shape in shapes:
        print("adjoint tile", shape, flush=True)
        Q.set_adjoint_tile(*shape)
        gr, ls = Q.adjoint(target, 1.0)
        b = Q.logproj_backward_saved(L, gp, 1.0)
        eg = float((gr - g0).abs().max()) / scale_g
        eb = float((b - b0).abs().max()) / scale_b
        el = abs(float(ls) - float(l0)) / float(l0)
        print(f"  grad {eg:.2e}  saved-backward {eb:.2e}  loss {el:.2e}")
        assert eg < 1e-5 and eb < 1e-5 and el < 1e-12
    Q.close()


def test_invalid_tile_shapes_are_refused(torch, case):
    from virtualcbct_amd._lib import VcbctError
    g, _, _, _ = case
    P = _projector(g)
    for bad in [(3, 1, 0, 0), (1, 3, 0, 0), (4, 4, 0, 0), (1, 4, 2, 0), (1, 4, 0, 2), (-1, 4, 0, 0), (2, 0, 0, 0)]:
        with pytest.raises(VcbctError, match="tile shape"):
            P.set_tile(*bad)
    P.set_tile(2, 4, 1, 1)
    P.set_tile(0)
    for bad in [(1, 2, 0, 0), (2, 4, 0, 0), (3, 1, 0, 0), (2, 2, 2, 0)]:
        with pytest.raises(VcbctError, match="tile shape"):
            P.set_adjoint_tile(*bad)
    P.set_adjoint_tile(2, 2, 1, 0)
    P.set_adjoint_tile(0)
    P.close()
