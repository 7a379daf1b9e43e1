"""Detector model on the device against the oracle's restatement (same generator, same samplers, FP64 on
both sides): counts equal pixel for pixel, barring a comparison that falls within an ulp of its threshold
(allowed: one pixel in 10^5, off by one).  At the full projection size (360 x 400 x 400), the standardised
counts (k - lambda) / sqrt(lambda) have mean 0 and variance 1."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as T
    if not T.cuda.is_available():
        pytest.skip("no GPU")
    return T


def _agree(got, want):
    diff = got != want
    frac = diff.mean()
    print(f"pixels differing from the oracle: {int(diff.sum())} of {diff.size}")
    assert frac <= 1e-5
    if diff.any():
        assert np.abs(got[diff] - want[diff]).max() <= 1.0


@pytest.mark.parametrize("seed", [0, 12345, (1 << 40) + 7])
def test_counts_match_oracle(torch, seed):
    from virtualcbct_amd import detector
    rng = np.random.default_rng(seed & 0xFFFF)
    n = 1 << 18
    # expectations from starved (1e-3 quanta) to bright (1e5), both samplers and their boundary at 10
    lam = np.exp(rng.uniform(np.log(1e-3), np.log(1e5), n))
    lam[:6] = [0.0, 9.999999, 10.0, 10.000001, 1e5, -1.0]
    prim = lam.astype(np.float32)
    want = O.detector_counts(prim, None, 1.0, seed, raw=True)
    got = detector.detector_counts(torch.from_numpy(prim).cuda(), I0=1.0, seed=seed, raw=True).cpu().numpy()
    _agree(got, want)


def test_scatter_normalised_and_in_place(torch):
    from virtualcbct_amd import detector
    rng = np.random.default_rng(8)
    shape = (3, 64, 200)
    prim = np.exp(-rng.uniform(0.0, 6.0, shape)).astype(np.float32)
    sc = (0.05 * rng.uniform(0.0, 1.0, shape)).astype(np.float32)
    I0 = 2.0e4
    want = O.detector_counts(prim, sc, I0, 77)
    dp, ds = torch.from_numpy(prim).cuda(), torch.from_numpy(sc).cuda()
    got = detector.detector_counts(dp, ds, I0=I0, seed=77).cpu().numpy()
    _agree(got, want)
    detector.detector_counts(dp, ds, I0=I0, seed=77, out=dp)  # in place
    np.testing.assert_array_equal(dp.cpu().numpy(), got)
    # a pixel's draw is keyed by its flat index, not by the launch: the first projection drawn alone (a
    # grid a third the size) gives the same counts as within the whole stack
    first = detector.detector_counts(torch.from_numpy(prim[:1].copy()).cuda(), torch.from_numpy(sc[:1].copy()).cuda(),
                                     I0=I0, seed=77).cpu().numpy()
    np.testing.assert_array_equal(first[0], got[0])
    with pytest.raises(Exception):
        detector.detector_counts(dp, I0=0.0)


def test_full_size_standardised_counts(torch):
    """c2's projection size (360 x 400 x 400 = 57.6 M pixels) at clinical and low dose."""
    from virtualcbct_amd import detector
    g = torch.Generator(device="cuda").manual_seed(3)
    prim = torch.rand((360, 400, 400), device="cuda", generator=g) * 0.9 + 0.05
    for I0 in (1.0e4, 30.0):
        k = detector.detector_counts(prim, I0=I0, seed=21, raw=True).double()
        lam = prim.double() * I0
        z = (k - lam) / lam.sqrt()
        m, v = float(z.mean()), float(z.var())
        print(f"I0 {I0:g}: standardised mean {m:.2e}, variance {v:.5f}")
        assert abs(m) < 5 / np.sqrt(z.numel())
        assert abs(v - 1) < 2e-3
        assert bool((k == k.round()).all()) and float(k.min()) >= 0
