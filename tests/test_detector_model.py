"""Detector model (scatter composition + quantum noise) oracle, pinned by the published known-answer vectors
of its generator and by the Poisson law it must sample.  No reference counterpart: the reference hands the
stage between its projector and FDK to other programs (projection file names in F/txtfile; the scatter
U-Net, Unet_scatter.py:349-423).  The GPU kernel is compared with this oracle in test_gpu_detector_model.py."""
import math

import numpy as np
import pytest

from oracle import oracle as O


@pytest.mark.parametrize("ctr,key,want", [
    # Random123 known-answer vectors for philox4x32 with 10 rounds
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_known_answers(ctr, key, want):
    assert O.philox4x32_10(ctr, key) == want


def test_log_factorial_against_lgamma():
    """The sampler's log k! (factorial table below 10, Stirling's series above) against the C library's lgamma."""
    for k in list(range(0, 40)) + [99, 1000, 12345, 10 ** 6, 3 * 10 ** 7]:
        assert O.log_factorial(k) == pytest.approx(math.lgamma(k + 1.0), rel=1e-11, abs=1e-14), k


def _counts(lam, n, seed=3):
    prim = np.full(n, lam, np.float32)
    return O.detector_counts(prim, None, 1.0, seed, raw=True).astype(np.float64)


@pytest.mark.parametrize("lam", [0.3, 2.5, 9.99, 10.0, 47.0, 1.0e4])
def test_poisson_moments(lam):
    """Mean and variance equal lambda (both samplers: inversion below 10, transformed rejection above)."""
    lam32 = float(np.float32(lam))
    n = 200000
    k = _counts(lam32, n)
    assert np.all(k == np.round(k)) and k.min() >= 0
    se = math.sqrt(lam32 / n)
    assert abs(k.mean() - lam32) < 5 * se
    # variance of the sample variance ~ (2 lam^2 + lam) / n
    assert abs(k.var() - lam32) < 5 * math.sqrt((2 * lam32 ** 2 + lam32) / n)


@pytest.mark.parametrize("lam", [3.0, 25.0])
def test_poisson_distribution(lam):
    """Chi-square of the sampled histogram against the Poisson pmf (a shifted or dropped term fails it)."""
    n = 200000
    k = _counts(lam, n, seed=11).astype(np.int64)
    kmax = int(lam + 6 * math.sqrt(lam)) + 2
    pmf = np.array([math.exp(-lam + j * math.log(lam) - math.lgamma(j + 1)) for j in range(kmax)])
    obs = np.bincount(np.minimum(k, kmax - 1), minlength=kmax)[:kmax].astype(float)
    exp_ = pmf * n
    exp_[-1] += n * (1 - pmf.sum())
    keep = exp_ > 20
    chi2 = float((((obs - exp_) ** 2) / exp_)[keep].sum())
    dof = int(keep.sum()) - 1
    # 5-sigma bound of the chi-square law
    assert chi2 < dof + 5 * math.sqrt(2 * dof), (chi2, dof)


def test_draw_depends_on_index_and_seed_only():
    rng = np.random.default_rng(2)
    prim = rng.uniform(0.0, 50.0, 3000).astype(np.float32)
    a = O.detector_counts(prim, None, 1.0, 99, raw=True)
    assert np.array_equal(a, O.detector_counts(prim, None, 1.0, 99, raw=True))
    assert not np.array_equal(a, O.detector_counts(prim, None, 1.0, 100, raw=True))
    for i in (0, 17, 2999):
        assert a[i] == O.poisson(float(np.float32(prim[i])), i, 99)
    # seeds above 32 bits use the key's second word
    hi = [O.poisson(30.0, i, 1 << 40) for i in range(50)]
    assert hi != [O.poisson(30.0, i, 0) for i in range(50)]


def test_scatter_adds_to_primary_and_normalisation():
    rng = np.random.default_rng(4)
    prim = rng.uniform(0.0, 1.0, 5000).astype(np.float32)
    sc = rng.uniform(0.0, 0.2, 5000).astype(np.float32)
    I0 = 1000.0
    raw = O.detector_counts(prim, sc, I0, 5, raw=True)
    lam = I0 * (prim.astype(np.float64) + sc.astype(np.float64))
    assert raw[0] == O.poisson(lam[0], 0, 5) and raw[4999] == O.poisson(lam[4999], 4999, 5)
    z = (raw - lam) / np.sqrt(lam)
    assert abs(z.mean()) < 0.1 and abs(z.std() - 1) < 0.05
    norm = O.detector_counts(prim, sc, I0, 5, raw=False)
    np.testing.assert_array_equal(norm, (np.maximum(raw.astype(np.float64), 1.0) / I0).astype(np.float32))


def test_zero_and_negative_expectation_give_zero_counts():
    prim = np.array([0.0, -1.0, np.nan, 1e-30], np.float32)
    raw = O.detector_counts(prim, None, 1.0, 1, raw=True)
    assert list(raw[:3]) == [0.0, 0.0, 0.0] and raw[3] == 0.0
    assert list(O.detector_counts(prim, None, 4.0, 1)) == [0.25] * 4
