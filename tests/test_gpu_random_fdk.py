"""Seeded random small FDK backprojections against the FP64 oracle (reconstructImage restated): random filtered
projections go straight to the backprojector, through the one-shot call and through an FDK plan in both its
forms (unstaged and LDS-staged, which must agree bit for bit).  The cases push the reference's edge cases:
panel shifts that move voxels onto the last column and past either edge, slice ranges beyond the first and
last rows, odd volume and panel sizes, partial 16 x 16 blocks and unsorted angle steps."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

SEEDS = list(range(16))


@pytest.fixture(scope="module")
def torch():
    import torch as T
    if not T.cuda.is_available():
        pytest.skip("no GPU")
    return T


def random_case(seed):
    rng = np.random.default_rng(9100 + seed)
    R = int(rng.integers(8, 49))
    Z = int(rng.integers(1, 10))
    n = int(rng.integers(3, 25))
    H = int(rng.integers(6, 61))
    W = int(rng.integers(6, 81))
    scale = float(rng.choice([0.05, 0.1, 0.2, 0.4]))
    step = float(rng.choice([0.05, 0.1, 0.25, 0.5]))
    zstart = int(rng.integers(-3, 2))
    ang = np.deg2rad(np.sort(rng.uniform(-180.0, 180.0, n)))
    # x shifts in the reference's unit (xshift W / 40 columns): small, or large enough to reach the panel edges
    xs = rng.uniform(-2.0, 2.0, n) if rng.random() < 0.5 else rng.uniform(-25.0, 25.0, n)
    sino = rng.uniform(-1.0, 1.0, (n, H, W)).astype(np.float32)
    xfactor = int(rng.choice([1, 1000]))
    return R, Z, n, H, W, scale, step, zstart, ang, xs, sino, xfactor


@pytest.mark.parametrize("seed", SEEDS)
def test_random_backprojection_matches_oracle(torch, monkeypatch, seed):
    from virtualcbct_amd import fdk
    R, Z, n, H, W, scale, step, zstart, ang, xs, sino, xfactor = random_case(seed)
    print(f"seed {seed}: R {R} Z {Z} angles {n} panel {H}x{W} scale {scale} step {step} zstart {zstart} "
          f"|xshift| <= {np.abs(xs).max():.1f} xfactor {xfactor}", flush=True)
    ref = O.reconstruct_image(R, Z, zstart, step, scale, sino, xfactor, ang, xs)
    scale_r = float(np.abs(ref).max())
    assert scale_r > 0
    d = torch.from_numpy(sino).cuda()
    rec = torch.zeros((Z, R, R), device="cuda")
    fdk.reconstruct_image_device(rec, d, zstart, step, scale, xfactor, ang, xs)
    e = float(np.abs(rec.cpu().numpy() - ref).max()) / scale_r
    print(f"  one-shot backprojection vs FP64 oracle: {e:.3e} of max")
    assert e < 1e-5
    recs = {}
    for lds in ("0", "1"):
        monkeypatch.setenv("VCBCT_BP_LDS", lds)
        plan = fdk.FdkPlan(H, W, n, R, Z, zstart, step, scale, xfactor, ang, xs, np.zeros(n))
        assert plan.bp_staged == (lds == "1")
        r = torch.zeros(plan.shape_recon, device="cuda")
        plan.backproject(d, r)
        torch.cuda.synchronize()
        recs[lds] = r
        plan.close()
    assert torch.equal(recs["0"], recs["1"])
    ep = float(np.abs(recs["0"].cpu().numpy() - ref).max()) / scale_r
    print(f"  plan backprojection vs FP64 oracle: {ep:.3e} of max (staged form equal bit for bit)")
    assert ep < 1e-5
