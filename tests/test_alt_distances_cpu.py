"""The Pearson and Anscombe SSD distances and the candidate-standardized
stage-1 policy on the CPU (blockmatching.py:29-87, 322-426).

Pins of the oracle: the null tables of the three distances against scipy's
binomial pmf (Pearson also against its closed form 1, 2 - 2/N); the
standardized distance against the reference's own property
(tests/test_blockmatching_internal_unit.py:232-266: over the binomial split
of a pooled count its scores have mean 0 and variance 1) and on patches
whose terms vanish; the plain distances by the reference's known answers in
tests/test_oracle_blockmatching.py.  Then the product's host tables and
thresholds against the oracle's, and the options the engine accepts."""

import math

import numpy as np
import pytest
from scipy.stats import binom

import oracle
from bmnd_amd import BM3DProfile, BM4DProfile, NoiseModel, blockmatching, phantoms
from bmnd_amd.engine import PoissonPipeline, check_supported
from bmnd_amd.enums import BlockMatchDistance, PoissonHTMatchPolicy
from bmnd_amd.profilevalidation import resolve_profile

CODES = {BlockMatchDistance.POISSON_DEVIANCE: oracle.DIST_DEVIANCE,
         BlockMatchDistance.PEARSON: oracle.DIST_PEARSON,
         BlockMatchDistance.ANSCOMBE_SSD: oracle.DIST_ANSCOMBE}


def _term(code, r, c):
    n = r + c
    if code == oracle.DIST_PEARSON:
        return (c - r) ** 2 / n
    if code == oracle.DIST_ANSCOMBE:
        return 0.5 * (2 * np.sqrt(c + 0.375) - 2 * np.sqrt(r + 0.375)) ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        return 2.0 * (np.where(r > 0, r * np.log(2.0 * r / n), 0.0)
                      + np.where(c > 0, c * np.log(2.0 * c / n), 0.0))


@pytest.mark.parametrize("code", [oracle.DIST_PEARSON, oracle.DIST_ANSCOMBE])
@pytest.mark.parametrize("n", [1, 2, 5, 40, 300])
def test_null_tables_match_scipy_binomial(code, n):
    means, variances = oracle.poisson_null_moments(300, code)
    r = np.arange(n + 1, dtype=np.float64)
    t = _term(code, r, n - r)
    p = binom.pmf(r, n, 0.5)
    e = float(np.sum(p * t))
    assert means[n] == pytest.approx(e, rel=1e-12, abs=1e-15)
    assert variances[n] == pytest.approx(float(np.sum(p * (t - e) ** 2)), rel=1e-9, abs=1e-14)
    if code == oracle.DIST_PEARSON:
        assert means[n] == 1.0 and variances[n] == pytest.approx(2.0 - 2.0 / n, rel=1e-15)


def _std_distance(ref, cand, code, tables, scale=1.0):
    m = np.ascontiguousarray(tables[0], np.float32)
    v = np.ascontiguousarray(tables[1], np.float32)
    r = np.ascontiguousarray(ref, np.float32)
    c = np.ascontiguousarray(cand, np.float32)
    return float(oracle.lib().oracle_standardized_distance(
        oracle._p(r), oracle._p(c), r.size, code, scale, oracle._p(m), oracle._p(v), len(m) - 1))


@pytest.mark.parametrize("code", [oracle.DIST_DEVIANCE, oracle.DIST_PEARSON, oracle.DIST_ANSCOMBE])
@pytest.mark.parametrize("pooled", [4, 9, 25])
def test_standardized_scores_have_unit_null_moments(code, pooled):
    """The reference's test (pooled count 4, table to 8), at more counts."""
    tables = oracle.poisson_null_moments(max(8, pooled), code)
    scores, probs = [], []
    for r in range(pooled + 1):
        scores.append(_std_distance([r], [pooled - r], code, tables))
        probs.append(math.comb(pooled, r) / 2 ** pooled)
    s, p = np.asarray(scores), np.asarray(probs)
    mean = float(np.sum(p * s))
    assert abs(mean) < 1e-6
    assert abs(float(np.sum(p * (s - mean) ** 2)) - 1.0) < 1e-6


@pytest.mark.parametrize("code", [oracle.DIST_DEVIANCE, oracle.DIST_PEARSON, oracle.DIST_ANSCOMBE])
def test_standardized_identical_and_empty_patches(code):
    tables = oracle.poisson_null_moments(16, code)
    m = np.float32(tables[0][6])
    v = np.float32(tables[1][6])
    # equal counts 3 in 8 voxels: no distance, the null sums of pooled count 6
    nm, nv = np.float32(0.0), np.float32(0.0)
    for _ in range(8):
        nm, nv = np.float32(nm + m), np.float32(nv + v)
    want = np.float32((0.0 - float(nm)) / float(np.sqrt(nv)))
    assert _std_distance(np.full(8, 3.0), np.full(8, 3.0), code, tables) == want
    # pooled counts above the table: chi-square(1) moments 1 and 2
    big = _std_distance(np.full(4, 40.0), np.full(4, 40.0), code, tables)
    assert big == np.float32(-4.0 / np.sqrt(np.float32(8.0)))
    # nothing positive: 0
    assert _std_distance(np.zeros(8), np.zeros(8), code, tables) == 0.0


def test_product_tables_equal_oracle():
    for dist, code in CODES.items():
        m, v = blockmatching.poisson_null_moment_table(dist, 200)
        om, ov = oracle.poisson_null_moments(200, code)
        assert np.array_equal(m, om) and np.array_equal(v, ov)
        f32 = blockmatching.poisson_null_moment_table_f32(dist, 200)
        assert np.array_equal(f32, np.concatenate([np.float32(om), np.float32(ov)]))
    with pytest.raises(ValueError, match="poisson_moment_max_count"):
        blockmatching.poisson_null_moment_table_f32(BlockMatchDistance.PEARSON, 0)


@pytest.mark.parametrize("dist", [BlockMatchDistance.PEARSON, BlockMatchDistance.ANSCOMBE_SSD])
def test_finite_count_thresholds_for_other_distances(dist):
    _, vol = phantoms.ellipsoid_volume((20, 22, 18), seed=83, peak=12.0, n_ellipsoids=4,
                                       axes_range=(3.0, 8.0))
    prof = BM4DProfile(noise_model=NoiseModel.POISSON)
    plan = PoissonPipeline(prof, vol.shape, "cpu").plans["ht"]
    got = blockmatching.compute_poisson_reference_thresholds(
        vol, plan.schedule, plan.block, 1.0, 2.9527, 0.1, 1.0, 512, distance=dist)
    params = oracle.default_params(3, vol.shape)
    params.distance = CODES[dist]
    want = oracle.poisson_reference_thresholds(vol, params, 2.9527, 0.1, 1.0)
    assert np.array_equal(got.view(np.int32), want.view(np.int32))


def test_engine_distance_and_policy_support():
    prof = BM3DProfile(noise_model=NoiseModel.POISSON)
    for dist in ("pearson", "anscombe_ssd"):
        for policy in ("fixed", "reference_finite_count", "candidate_standardized"):
            prof.blockmatch_distance = dist
            prof.poisson_ht_match_policy = policy
            check_supported(prof, resolve_profile(prof, (32, 32)))
    # SSD under the Poisson model: the fixed policy only (the reference's
    # validation refuses the others, profilevalidation.py:414-425)
    prof.blockmatch_distance = "ssd"
    prof.poisson_ht_match_policy = PoissonHTMatchPolicy.FIXED
    check_supported(prof, resolve_profile(prof, (32, 32)))
    pipe = PoissonPipeline(prof, (32, 32), "cpu")
    assert pipe.blockmatch_mode(None, "ht") == (11, 0)  # gamma 3: the Poisson correction
    assert pipe.blockmatch_mode(None, "wiener") == (10, 0)
    assert not pipe.graph_supported()
    prof.gamma = 0.0
    assert PoissonPipeline(prof, (32, 32), "cpu").blockmatch_mode(None, "ht") == (10, 0)
    prof.poisson_ht_match_policy = PoissonHTMatchPolicy.REFERENCE_FINITE_COUNT
    with pytest.raises(ValueError, match="requires"):
        resolve_profile(prof, (32, 32))


@pytest.mark.parametrize("dist,policy,ht,wi", [
    ("pearson", "fixed", 5, 5), ("anscombe_ssd", "fixed", 6, 6),
    ("poisson_deviance", "candidate_standardized", 7, None),
    ("pearson", "candidate_standardized", 8, 5), ("anscombe_ssd", "candidate_standardized", 9, 6)])
def test_engine_modes(dist, policy, ht, wi):
    """Standardization applies to stage 1 only (the reference's
    tests/test_bmnd_pipeline_unit.py:838-870)."""
    prof = BM3DProfile(noise_model=NoiseModel.POISSON)
    prof.blockmatch_distance = dist
    prof.poisson_ht_match_policy = policy
    pipe = PoissonPipeline(prof, (32, 32), "cpu")
    assert pipe.blockmatch_mode(None, "ht") == (ht, 0)
    if wi is not None:
        assert pipe.blockmatch_mode(None, "wiener") == (wi, 0)


# ---- Wiener gain modes (wiener.py:24-53, 78-95) ------------------------------

@pytest.mark.parametrize("mu,ref,checks", [
    # tests/compat/test_component_semantics.py:54-76 (mu = 1)
    (1.0, [0.0, 0.1, 0.4, 1.5], (("classic", 0.0), ("variance-scaled", 1.0))),
    # tests/compat/test_component_semantics.py:79-108 (mu = 2)
    (2.0, [0.0, 0.3, 0.6, 1.5], (("classic", 0.0), ("noise-floor", 1.0),
                                 ("variance-scaled", 2.0)))])
def test_gain_modes_known_answers(mu, ref, checks):
    """The reference's closed forms: signal = max(r^2 - s sigma^2, 0) with
    s = 0 (classic), 1 (noise-floor), mu (variance-scaled); gain = signal /
    (signal + mu sigma^2)."""
    r = np.asarray(ref, np.float32)
    sigma = np.full(4, 0.2, np.float32)
    for mode, sub in checks:
        coef, gain = oracle.wiener_shrink(np.ones(4, np.float32), r, sigma, mu, mode)
        signal = np.maximum(r.astype(np.float64) ** 2 - sub * 0.04, 0.0)
        np.testing.assert_allclose(gain, signal / (signal + mu * 0.04), rtol=1e-6)
        assert np.array_equal(coef, gain)


def test_gain_modes_are_ordered():
    """Classic >= noise-floor >= variance-scaled at mu >= 1 (the reference's
    'sits between' property)."""
    rng = np.random.default_rng(5)
    r = rng.normal(0.0, 1.0, 4096).astype(np.float32)
    sigma = rng.uniform(0.05, 1.0, 4096).astype(np.float32)
    g = {m: oracle.wiener_shrink(np.ones_like(r), r, sigma, 2.0, m)[1]
         for m in ("classic", "noise-floor", "variance-scaled")}
    assert np.all(g["classic"] >= g["noise-floor"])
    assert np.all(g["noise-floor"] >= g["variance-scaled"])


@pytest.mark.parametrize("mode,arg", [("variance-scaled", 1), ("classic", 2), ("noise-floor", 3)])
def test_engine_gain_argument(mode, arg):
    prof = BM4DProfile(noise_model=NoiseModel.POISSON)
    prof.poisson_wiener_gain_mode = mode
    check_supported(prof, resolve_profile(prof, (16, 16, 16)))
    assert PoissonPipeline(prof, (16, 16, 16), "cpu").wiener_arg == arg


@pytest.mark.parametrize("dist,policy,ht,wi", [
    ("pearson", "fixed", 5, 5), ("anscombe_ssd", "candidate_standardized", 9, 6),
    ("poisson_deviance", "candidate_standardized", 7, None)])
def test_batch_modes_follow_the_single_pipeline(dist, policy, ht, wi):
    """A batch matches every input in the mode the single-input pipeline
    picks; these options need no look at the data."""
    from bmnd_amd.engine import BatchPipeline
    prof = BM3DProfile(noise_model=NoiseModel.POISSON)
    prof.blockmatch_distance = dist
    prof.poisson_ht_match_policy = policy
    bp = BatchPipeline(prof, [(32, 40), (20, 64), (48, 48)], "cpu")
    assert bp.input_modes(None, "ht") == [(ht, 0)] * 3
    if wi is not None:
        assert bp.input_modes(None, "wiener") == [(wi, 0)] * 3
