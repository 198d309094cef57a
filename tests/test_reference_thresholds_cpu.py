"""Per-reference stage-1 thresholds of poisson_ht_match_policy=
"reference_finite_count" (blockmatching.py:29-148, bmndalgo.py:269-284), on
the CPU.

The oracle's null-moment table is pinned by closed forms (N = 1 and 2), by
an independent computation from scipy's binomial pmf, and by the
chi-square(1) limit (mean 1, variance 2) that the reference substitutes above
the table.  The oracle's per-reference loop is pinned on constant patches,
whose thresholds reduce to a single table entry.  The product's vectorised
host function must then equal the oracle's loop bit for bit."""

import math

import numpy as np
import pytest
from scipy.stats import binom

import oracle
from bmnd_amd import BM3DProfile, BM4DProfile, NoiseModel, blockmatching, phantoms
from bmnd_amd.engine import PoissonPipeline, check_supported
from bmnd_amd.enums import PoissonHTMatchPolicy
from bmnd_amd.profilevalidation import resolve_profile


@pytest.fixture(scope="module")
def table():
    return oracle.poisson_null_moments(512)


def test_null_moments_closed_forms(table):
    means, variances = table
    ln2 = math.log(2.0)
    assert means[0] == 0.0 and variances[0] == 0.0
    # N = 1: both splits give 2 log 2
    assert means[1] == pytest.approx(2 * ln2, rel=1e-15) and variances[1] == 0.0
    # N = 2: 4 log 2 with probability 1/2, 0 otherwise
    assert means[2] == pytest.approx(2 * ln2, rel=1e-15)
    assert variances[2] == pytest.approx(4 * ln2 * ln2, rel=1e-14)


@pytest.mark.parametrize("n", [3, 7, 20, 100, 300, 512])
def test_null_moments_match_scipy_binomial(table, n):
    means, variances = table
    r = np.arange(n + 1, dtype=np.float64)
    c = n - r
    with np.errstate(divide="ignore", invalid="ignore"):
        t = 2.0 * (np.where(r > 0, r * np.log(2.0 * r / n), 0.0)
                   + np.where(c > 0, c * np.log(2.0 * c / n), 0.0))
    p = binom.pmf(r, n, 0.5)
    e = float(np.sum(p * t))
    var = float(np.sum(p * (t - e) ** 2))
    assert means[n] == pytest.approx(e, rel=1e-12)
    assert variances[n] == pytest.approx(var, rel=1e-9)


def test_null_moments_approach_chi_square(table):
    means, variances = table
    assert abs(means[512] - 1.0) < 2e-3 and abs(variances[512] - 2.0) < 1e-2


def test_product_table_equals_oracle(table):
    m, v = blockmatching.poisson_null_moment_table(max_count=512)
    assert np.array_equal(m, np.asarray(table[0])) and np.array_equal(v, np.asarray(table[1]))


@pytest.mark.parametrize("level,above", [(3.0, False), (0.25, False), (700.0, True)])
def test_constant_patches_reduce_to_one_entry(table, level, above):
    means, variances = table
    vol = np.full((20, 22, 18), level, np.float32)
    params = oracle.default_params(3, vol.shape)
    lam, beta, power = 2.9527, 0.3, 0.75
    thr = oracle.poisson_reference_thresholds(vol, params, lam, beta, power)
    P = 64
    n = int(np.rint(2.0 * level))
    mu, var = (1.0, 2.0) if above else (means[n], variances[n])
    want = np.float32(mu + lam * math.sqrt(var / P) + beta * level ** power)
    assert thr.size > 0 and np.all(thr == want)


def _inputs():
    _, v3 = phantoms.ellipsoid_volume((26, 24, 22), seed=81, peak=12.0, n_ellipsoids=5,
                                      axes_range=(3.0, 9.0))
    _, v2 = phantoms.blob_image(48, seed=82, peak=30.0)
    return v3, v2


@pytest.mark.parametrize("scale,beta,power,max_count", [
    (1.0, 0.0, 1.0, 512), (1.0, 0.4, 0.5, 512), (2.5, 0.1, 1.5, 512), (1.0, 0.2, 1.0, 16)])
def test_vectorised_host_thresholds_equal_oracle_loop(scale, beta, power, max_count):
    for vol in _inputs():
        prof = (BM4DProfile if vol.ndim == 3 else BM3DProfile)(noise_model=NoiseModel.POISSON)
        pipe = PoissonPipeline(prof, vol.shape, "cpu")
        plan = pipe.plans["ht"]
        got = blockmatching.compute_poisson_reference_thresholds(
            vol, plan.schedule, plan.block, scale, prof.ht_match_threshold, beta, power,
            max_count)
        params = oracle.default_params(vol.ndim, vol.shape)
        params.count_scale = scale
        want = oracle.poisson_reference_thresholds(vol, params, prof.ht_match_threshold, beta,
                                                   power, max_count)
        assert got.dtype == np.float32 and got.shape == want.shape
        assert np.array_equal(got.view(np.int32), want.view(np.int32))


def test_threshold_option_errors():
    vol = np.ones((12, 12), np.float32)
    prof = BM3DProfile(noise_model=NoiseModel.POISSON)
    plan = PoissonPipeline(prof, vol.shape, "cpu").plans["ht"]
    f = blockmatching.compute_poisson_reference_thresholds
    with pytest.raises(ValueError, match="ht_match_threshold"):
        f(vol, plan.schedule, plan.block, 1.0, -1.0, 0.0, 1.0)
    with pytest.raises(ValueError, match="structure_beta"):
        f(vol, plan.schedule, plan.block, 1.0, 1.0, float("nan"), 1.0)
    with pytest.raises(ValueError, match="intensity_power"):
        f(vol, plan.schedule, plan.block, 1.0, 1.0, 0.0, 0.0)


def test_policy_support():
    prof = BM4DProfile(noise_model=NoiseModel.POISSON)
    prof.poisson_ht_match_policy = PoissonHTMatchPolicy.REFERENCE_FINITE_COUNT
    check_supported(prof, resolve_profile(prof, (16, 16, 16)))
    prof.poisson_ht_match_policy = PoissonHTMatchPolicy.CANDIDATE_STANDARDIZED
    check_supported(prof, resolve_profile(prof, (16, 16, 16)))
    prof.poisson_group_mass_conservation = "ht"
    check_supported(prof, resolve_profile(prof, (16, 16, 16)))
    prof.wiener_weight_domain = "windowed_synthesis"
    prof.wiener_weight_scope = "group"
    check_supported(prof, resolve_profile(prof, (16, 16, 16)))  # the exact group risk
