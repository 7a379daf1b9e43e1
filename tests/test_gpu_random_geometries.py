"""Seeded random small geometries against the FP64 oracle: the intensity projection of a u16 volume, the log
projection and the adjoint of FP32 maps.  The cases vary what the fixed-geometry tests hold constant: voxel
pitches on and off the whole-cm plane grid, slice thickness against detector-row spacing (from rows several
slices apart to several rows per slice, where runs of lanes share a voxel in the scatter), row counts across
the 64-lane wave boundary, detector shifts, isocentre offsets, angles near the 45 / 135 degree diagonals, and
rank-2 and full-rank attenuation tables (the spectral channel volume and the six maps)."""
import numpy as np
import pytest

from oracle import oracle as O
from virtualcbct_amd import phantoms

pytestmark = pytest.mark.gpu

SEEDS = list(range(24))


@pytest.fixture(scope="module")
def torch():
    import torch as T
    if not T.cuda.is_available():
        pytest.skip("no GPU")
    return T


@pytest.fixture(scope="module")
def rank2(tmp_path_factory):
    return phantoms.synthetic_spectrum(str(tmp_path_factory.mktemp("spec")), 90)


def random_case(seed):
    rng = np.random.default_rng(7000 + seed)
    N = int(rng.integers(12, 41))
    s = float(rng.choice([0.25, 0.3, 0.37, 0.5, 0.8]))
    Nz = int(rng.integers(3, 13))
    t = float(rng.choice([0.1, 0.2, 0.33, 0.5, 1.0]))
    H = int(rng.choice([7, 33, 64, 65, 100]))
    # detector pitch at the isocentre (rows and columns alike: 0.0325 cm x pixnum / H) from 0.13 to 3 slices, so
    # that from 8 rows per slice to one row every 3 slices; columns spanning about the volume's width
    sp = t * float(rng.choice([0.13, 0.3, 0.7, 1.5, 3.0]))
    pixnum = sp * H / 0.032509752
    W = int(np.clip(np.ceil(N * s / sp * rng.uniform(0.6, 1.2)), 5, 64))
    nA = int(rng.integers(2, 6))
    base = rng.uniform(-179.0, 179.0, nA)
    diag = rng.choice([45.0, 135.0, -45.0, -135.0], nA) + rng.uniform(-0.3, 0.3, nA)
    ang = np.deg2rad(np.where(rng.random(nA) < 0.4, diag, base) + 1e-5)
    # the column centre is -H/2 - k xshift with k = 10 / (0.05 pixnum / H) (P/misc.cpp:25-34 uses H, not W):
    # xshift = (W - H) / (2k) centres the W columns; a random shift of up to a fifth of the panel moves them, and
    # the rows, off centre
    k = 10.0 / (0.05 * (pixnum / H))
    xs = ((W - H) / 2.0 + rng.uniform(-0.2, 0.2, nA) * W) / k
    ys = rng.uniform(-0.1, 0.1, nA) * H / k
    cs, sn, xc, yc = O.trig_table_proj(W, H, ang, xs, ys, det_pixel_num=pixnum)
    g = O.Geometry(N=N, s=s, Nz=Nz, t=t, H=H, W=W, cos_t=cs, sin_t=sn, xc=xc, yc=yc,
                   x_reg=float(rng.uniform(-0.3, 0.3)), y_reg=float(rng.uniform(-0.3, 0.3)),
                   z_reg=float(rng.uniform(-0.4, 0.4)), det_pixel_num=pixnum)
    vol = rng.integers(0, 2000, (Nz, 6, N, N)).astype(np.uint16)
    vol[rng.random(vol.shape) < 0.3] = 0
    full_rank = bool(rng.random() < 0.5)
    NE = int(rng.integers(3, 12))
    atten = rng.uniform(0.05, 0.5, (6, NE)).astype(np.float32)
    pdf = rng.uniform(0.1, 1.0, NE).astype(np.float32)
    return g, vol, full_rank, (atten, pdf), rng


def projector(g):
    from virtualcbct_amd.projector import Projector
    return Projector(g.N, g.s, g.Nz, g.t, g.H, g.W, g.n_ang, g.cos_t, g.sin_t, g.xc, g.yc, g.x_reg, g.y_reg, g.z_reg,
                     g.det_pitch, g.det_pitch_iso, g.sad, g.det_pixel_num, g.nmat, 0)


def describe(seed, g, full_rank):
    return (f"seed {seed}: N {g.N} s {g.s} Nz {g.Nz} t {g.t} H {g.H} W {g.W} angles {g.n_ang} "
            f"pixnum {g.det_pixel_num} table {'full rank' if full_rank else 'rank 2'}")


@pytest.mark.parametrize("seed", SEEDS)
def test_random_geometry_forward_logproj_adjoint_match_oracle(torch, rank2, seed):
    g, vol, full_rank, table, rng = random_case(seed)
    atten, pdf = table if full_rank else rank2
    print(describe(seed, g, full_rank), flush=True)
    P = projector(g)
    P.set_volume(vol)
    P.set_spectrum(atten, pdf)
    out = P.forward(1000.0).cpu().numpy()
    ref = O.forward_collapsed(g, vol, atten, pdf, 1000.0)
    e = float(np.max(np.abs(out.astype(np.float64) - ref) / np.maximum(np.abs(ref), 1e-30)))
    print(f"  intensity vs FP64 oracle: {e:.3e} (forward channels {P.forward_channels})")
    assert e < 1e-4
    P.close()

    maps = np.ascontiguousarray(vol.transpose(1, 0, 2, 3)).astype(np.float32) / 1000.0
    vz = np.ascontiguousarray(maps.transpose(1, 0, 2, 3))
    Q = projector(g)
    Q.set_volume(torch.from_numpy(maps).cuda())
    Q.set_spectrum(atten, pdf)
    p = Q.logproj(1.0).cpu().numpy()
    po = O.logproj(g, vz, atten, pdf, 1.0)
    scale_p = max(float(np.abs(po).max()), 1e-30)
    ep = float(np.abs(p - po).max()) / scale_p
    print(f"  log projection vs FP64 oracle: {ep:.3e} of max")
    assert ep < 1e-5
    target = (po + rng.normal(0, 0.05, po.shape)).astype(np.float32)
    grad, loss = Q.adjoint(torch.from_numpy(target).cuda(), 1.0)
    go, lo = O.adjoint(g, vz, atten, pdf, 1.0, target.astype(np.float64))
    go = go.transpose(1, 0, 2, 3)
    scale_g = float(np.abs(go).max())
    nch = Q.adjoint_channels
    Q.close()
    if scale_g == 0.0:  # no ray of this geometry reaches the volume
        assert float(grad.abs().max()) == 0.0
        return
    eg = float(np.abs(grad.cpu().numpy() - go).max()) / scale_g
    print(f"  adjoint vs FP64 oracle: {eg:.3e} of max|grad| ({nch} channels); loss {float(loss):.6e} "
          f"vs {lo:.6e}")
    assert eg < 1e-4 and float(loss) == pytest.approx(lo, rel=1e-4)
