"""CPU checks of the sigma-map pieces (bmnd(..., return_sigma_map=True)).

numpy's float32 summation order (np.sum / np.mean), which the reference's
Wiener group sigma (bmndalgo.py:874-875) and global sigma
(sigmaestimation.py:62-75) inherit, is pinned against numpy itself, both in
the oracle and in the library's host evaluation of the summation plans its
kernels use.  The oracle's global sigma is pinned against numpy (the
reference's branches: all positive, positive subset, sparse trimmed mean)
and scipy's trimmed mean.  The oracle's sigma maps are pinned by closed forms
and invariants: a constant volume with k = 0 has maps sqrt(c); with k = 0
every map value is a weighted mean of coefficient stds of an orthonormal or
unit-gain transform, so it lies within the square roots of the stage's
variance range; the estimate is the same with and without maps.  No
reference vectors pin the maps voxel by voxel (the reference cannot run
here).
"""

import numpy as np
import pytest
from scipy import stats

import oracle
from bmnd_amd import phantoms


def _lengths(rng):
    small = list(range(1, 300))
    large = [int(x) for x in rng.integers(300, 120_000, 60)]
    edges = [8191, 8192, 8193, 16384, 16385, 4000, 2048, 3 * 8192 + 129]
    return small + large + edges


def test_oracle_np_sum_and_mean_match_numpy():
    rng = np.random.default_rng(21)
    for n in _lengths(rng):
        a = (rng.random(n) * rng.choice([1e-3, 1.0, 1e4])).astype(np.float32)
        assert oracle.np_sum_f32(a) == np.sum(a), n
        assert oracle.np_mean_f32(a) == np.mean(a), n
    vol = (rng.random((64, 64, 40)) * 7.0).astype(np.float32)
    assert oracle.np_mean_f32(vol) == np.mean(vol)
    assert oracle.np_mean_f32(vol) == np.mean(vol, dtype=np.float32)


def test_library_host_plan_sum_matches_numpy():
    """The kernels' summation plans, evaluated on the host by the library."""
    from bmnd_amd import _lib
    L = _lib.lib()
    rng = np.random.default_rng(22)
    for n in _lengths(rng):
        a = (rng.random(n) * rng.choice([1e-3, 1.0, 1e4])).astype(np.float32)
        got = np.float32(L.bmnd_hip_np_sum_host(a.ctypes.data, n))
        assert got == np.sum(a), n


def _sqrt_max(mean, eps):
    return float(np.sqrt(max(float(mean), eps)))


@pytest.mark.parametrize("floor", [1e-10, 0.25])
def test_global_sigma_all_positive_is_numpy_mean(floor):
    rng = np.random.default_rng(23)
    var = np.maximum(rng.poisson(4.0, (40, 37, 29)).astype(np.float32), np.float32(floor))
    eps = max(floor, 1e-10)
    assert oracle.poisson_global_sigma(var, floor) == _sqrt_max(np.mean(var, dtype=np.float32), eps)


def test_global_sigma_positive_subset():
    """A zero floor leaves zeros in the variance volume; with more than a
    tenth of the values positive the mean runs over the positive ones."""
    rng = np.random.default_rng(24)
    var = rng.poisson(0.7, (50, 60, 31)).astype(np.float32)
    pos = var[var > 0]
    assert 0.1 < pos.size / var.size < 1.0
    assert oracle.poisson_global_sigma(var, 0.0) == _sqrt_max(np.mean(pos), 1e-10)


@pytest.mark.parametrize("seed", [25, 26])
def test_global_sigma_sparse_trimmed_mean(seed):
    """Fewer than a tenth positive: the 10%-trimmed mean of the sorted
    positive values (scipy's trim_mean cuts the same counts)."""
    rng = np.random.default_rng(seed)
    var = np.zeros((40, 50, 33), np.float32)
    idx = rng.choice(var.size, size=int(0.04 * var.size), replace=False)
    var.ravel()[idx] = rng.gamma(2.0, 3.0, idx.size).astype(np.float32)
    pos = var[var > 0]
    trim = max(int(0.1 * pos.size), 1)
    got = oracle.poisson_global_sigma(var, 0.0)
    assert got == _sqrt_max(np.mean(np.sort(pos)[trim:-trim]), 1e-10)
    assert got == pytest.approx(float(np.sqrt(stats.trim_mean(pos, 0.1))), rel=1e-6)
    # the untrimmed mean is measurably different on these values
    assert abs(got - _sqrt_max(np.mean(pos), 1e-10)) > 1e-4


def test_global_sigma_degenerate_cases():
    eps = 1e-10
    assert oracle.poisson_global_sigma(np.zeros(1000, np.float32), 0.0) == np.sqrt(eps)
    two = np.zeros(1000, np.float32)
    two[[3, 700]] = [4.0, 9.0]  # fewer than 3 positive values: plain mean
    assert oracle.poisson_global_sigma(two, 0.0) == _sqrt_max(np.float32(6.5), eps)
    withinf = np.full(100, 2.0, np.float32)
    withinf[5] = np.inf  # non-finite values are left out
    assert oracle.poisson_global_sigma(withinf, 0.0) == _sqrt_max(2.0, eps)
    tiny = np.full(50, 1e-12, np.float32)  # a mean below eps
    assert oracle.poisson_global_sigma(tiny, 1e-10) == np.sqrt(1e-10)


@pytest.mark.parametrize("c", [3.0, 50.0])
def test_oracle_sigma_maps_constant_volume(c):
    v = np.full((16, 18, 17), c, np.float32)
    p = oracle.default_params(3, v.shape)
    p.k = 0
    out, maps = oracle.bmnd_poisson_sigma_maps(v, p)
    assert maps["global"] == np.sqrt(c)
    for stage in ("ht", "wiener"):
        np.testing.assert_allclose(maps[stage], np.sqrt(c), rtol=3e-5)


def test_oracle_sigma_maps_lie_in_the_stage_variance_range():
    _, noisy = phantoms.ellipsoid_volume((20, 22, 18), seed=1, peak=10.0, n_ellipsoids=4,
                                         axes_range=(3.0, 8.0))
    p = oracle.default_params(3, noisy.shape)
    p.k = 0
    out, pilot = oracle.bmnd_poisson(noisy, p)
    out2, maps = oracle.bmnd_poisson_sigma_maps(noisy, p)
    assert np.array_equal(out.view(np.int32), out2.view(np.int32))
    for stage, src in (("ht", noisy), ("wiener", pilot)):
        var = np.maximum(src, np.float32(1e-10))
        lo, hi = np.sqrt(var.min()), np.sqrt(var.max())
        assert maps[stage].min() >= lo * (1 - 1e-5) and maps[stage].max() <= hi * (1 + 1e-5)
        assert np.all(np.isfinite(maps[stage]))


def test_oracle_sigma_maps_keep_the_estimate_and_contract():
    _, img = phantoms.blob_image(40, seed=3, peak=10.0)
    ref, _ = oracle.bmnd_poisson(img)
    out, maps = oracle.bmnd_poisson_sigma_maps(img)
    assert np.array_equal(ref.view(np.int32), out.view(np.int32))
    assert set(maps) == {"global", "ht", "wiener"}
    assert isinstance(maps["global"], float)
    assert maps["global"] == _sqrt_max(np.mean(np.maximum(img, np.float32(1e-10))), 1e-10)
    for stage in ("ht", "wiener"):
        assert maps[stage].shape == img.shape and maps[stage].dtype == np.float32
        assert np.all(maps[stage] > 0)
    z = np.zeros((8, 8), np.float32)
    zo, zm = oracle.bmnd_poisson_sigma_maps(z)
    assert set(zm) == {"ht", "wiener"} and not np.any(zm["ht"]) and not np.any(zm["wiener"])
