"""The forward over the spectral channel volume.  When the attenuation table has a rank-r form (r <= 4) that
reproduces every energy column to 1e-6 of its largest entry, the intensity and log projections and the adjoint's
coefficient kernels traverse c_k = sum_m B_mk rho_m (2 or 4 FP32 per voxel) instead of the six maps.  Checked
here: which tables select it; projections and gradients through it against the FP64 oracle (which sums over the
full material table); with two materials (no basis: the channels are the maps) the kept line integrals of a u16
volume equal its material line integrals bit for bit; the channel volume follows every change of volume and
spectrum."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from virtualcbct_amd import phantoms

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as T
    if not T.cuda.is_available():
        pytest.skip("no GPU")
    return T


@pytest.fixture(scope="module")
def rank2(tmp_path_factory):
    """The synthetic workloads' table (photoelectric + Compton terms per element: rank 2)."""
    return phantoms.synthetic_spectrum(str(tmp_path_factory.mktemp("spec")), 100)


def geometry(nmat=6, N=32, Nz=10, H=40, W=36):
    ang = np.deg2rad(np.array([-170.0001, -100.0001, -45.0, 9.99999, 79.99999, 135.0]))
    cs, sn, xc, yc = O.trig_table_proj(W, H, ang, np.full(6, 0.02), np.zeros(6))
    return O.Geometry(N=N, s=0.5, Nz=Nz, t=0.5, H=H, W=W, cos_t=cs, sin_t=sn, xc=xc, yc=yc, z_reg=0.1, nmat=nmat)


def projector(g):
    from virtualcbct_amd.projector import Projector
    return Projector(g.N, g.s, g.Nz, g.t, g.H, g.W, g.n_ang, g.cos_t, g.sin_t, g.xc, g.yc, g.x_reg, g.y_reg, g.z_reg,
                     g.det_pitch, g.det_pitch_iso, g.sad, g.det_pixel_num, g.nmat, 0)


def rel(a, b):
    return float(np.max(np.abs(a.astype(np.float64) - b) / np.maximum(np.abs(b), 1e-30)))


def test_rank2_table_projects_through_two_channels(torch, rank2):
    atten, pdf = rank2
    g = geometry()
    vol = phantoms.ellipsoid_phantom(g.N, g.Nz, g.s, g.t, seed=31, n_ell=8, n_labels=5)
    P = projector(g)
    P.set_volume(vol)
    P.set_spectrum(atten, pdf)
    assert P.adjoint_channels == 2 and P.forward_channels == 2
    out = P.forward(1000.0).cpu().numpy()
    ref = O.forward_collapsed(g, vol, atten, pdf, 1000.0)
    e = rel(out, ref)
    print(f"intensity through 2 channels vs FP64 oracle: {e:.3e}")
    assert e < 1e-5
    # FP32 maps: log projection, the kept channel integrals, the gradient and its saved-L form
    maps = np.ascontiguousarray(vol.transpose(1, 0, 2, 3)).astype(np.float32) / 1000.0
    Q = projector(g)
    Q.set_volume(torch.from_numpy(maps).cuda())
    Q.set_spectrum(atten, pdf)
    p = Q.logproj(1.0).cpu().numpy()
    vz = np.ascontiguousarray(maps.transpose(1, 0, 2, 3))
    po = O.logproj(g, vz, atten, pdf, 1.0)
    ep = float(np.abs(p - po).max() / np.abs(po).max())
    print(f"log projection through 2 channels vs FP64 oracle: {ep:.3e} of max")
    assert ep < 1e-5
    p2, L = Q.logproj_save(1.0)
    assert torch.equal(torch.from_numpy(p).cuda(), p2)
    target = po * 0.97 + 0.01
    grad, loss = Q.adjoint(torch.from_numpy(target.astype(np.float32)).cuda(), 1.0)
    go, lo = O.adjoint(g, vz, atten, pdf, 1.0, target.astype(np.float32).astype(np.float64))
    go = go.transpose(1, 0, 2, 3)
    eg = float(np.abs(grad.cpu().numpy() - go).max() / np.abs(go).max())
    print(f"gradient through 2 channels vs FP64 oracle: {eg:.3e} of max; loss {float(loss):.6e} vs {lo:.6e}")
    assert eg < 1e-4 and float(loss) == pytest.approx(lo, rel=1e-5)
    gp = torch.from_numpy((2 * (p - target)).astype(np.float32)).cuda()
    g2 = Q.logproj_backward_saved(L, gp, 1.0).cpu().numpy()
    eg2 = float(np.abs(g2 - go).max() / np.abs(go).max())
    print(f"gradient from the kept channel integrals vs FP64 oracle: {eg2:.3e}")
    assert eg2 < 1e-4
    P.close()
    Q.close()


def test_full_rank_table_keeps_the_maps(torch):
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "reproj_small.npz"))
    g = geometry()
    P = projector(g)
    P.set_spectrum(d["atten"], d["pdf"])  # 6 x 5 random entries: rank 5, more channels than the forward takes
    assert P.forward_channels == 0 and P.adjoint_channels == 5
    P.close()


def test_two_materials_channels_are_the_maps(torch):
    """num_material = 2 with a random (full-rank) table: no basis, the channel volume holds the maps as FP32.
    From a u16 volume every product and sum is the u16 path's, in its order: the kept line integrals equal
    line_integrals() bit for bit.  From FP32 maps the sums run in another order (the FP32 maps' accumulator
    does not pipeline its first gather): equal to FP32 rounding."""
    rng = np.random.default_rng(5)
    g = geometry(nmat=2)
    NE = 9
    atten = rng.uniform(0.1, 0.5, (2, NE)).astype(np.float32)
    pdf = rng.uniform(0.2, 1.0, NE).astype(np.float32)
    vol = rng.integers(0, 1500, (g.Nz, 2, g.N, g.N)).astype(np.uint16)
    P = projector(g)
    P.set_volume(vol)
    P.set_spectrum(atten, pdf)
    assert P.forward_channels == 2 and P.adjoint_channels == 2
    _, L = P.logproj_save(1000.0)
    Lm = P.line_integrals().reshape(-1, 2).t().contiguous()
    assert torch.equal(L, Lm)
    out = P.forward(1000.0).cpu().numpy()
    assert rel(out, O.forward_collapsed(g, vol, atten, pdf, 1000.0)) < 1e-5
    maps = np.ascontiguousarray(vol.transpose(1, 0, 2, 3)).astype(np.float32) / 1000.0
    Q = projector(g)
    Q.set_volume(torch.from_numpy(maps).cuda())
    Q.set_spectrum(atten, pdf)
    _, Lq = Q.logproj_save(1.0)
    Lqm = Q.line_integrals().reshape(-1, 2).t().contiguous()
    scale = float(Lqm.abs().max())
    assert scale > 0 and float((Lq - Lqm).abs().max()) <= 1e-6 * scale
    p = Q.logproj(1.0).cpu().numpy()
    po = O.logproj(g, np.ascontiguousarray(maps.transpose(1, 0, 2, 3)), atten, pdf, 1.0)
    assert float(np.abs(p - po).max() / np.abs(po).max()) < 1e-5
    P.close()
    Q.close()


def test_channel_volume_follows_volume_and_spectrum(torch, rank2):
    atten, pdf = rank2
    g = geometry()
    va = phantoms.ellipsoid_phantom(g.N, g.Nz, g.s, g.t, seed=41, n_ell=8, n_labels=5)
    vb = phantoms.ellipsoid_phantom(g.N, g.Nz, g.s, g.t, seed=42, n_ell=8, n_labels=5)
    P = projector(g)
    P.set_spectrum(atten, pdf)
    P.set_volume(va)
    a = P.forward(1000.0).clone()
    P.set_volume(vb)
    b = P.forward(1000.0).clone()
    Q = projector(g)
    Q.set_spectrum(atten, pdf)
    Q.set_volume(vb)
    assert torch.equal(b, Q.forward(1000.0)) and not torch.equal(a, b)
    # spectrum change to a full-rank table: back to the maps; and back again
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "reproj_small.npz"))
    P.set_spectrum(d["atten"], d["pdf"])
    assert P.forward_channels == 0
    c = P.forward(1000.0).cpu().numpy()
    assert rel(c, O.forward_collapsed(g, vb, d["atten"], d["pdf"], 1000.0)) < 1e-5
    P.set_spectrum(atten, pdf)
    assert P.forward_channels == 2 and torch.equal(P.forward(1000.0), b)
    P.close()
    Q.close()
