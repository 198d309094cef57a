"""GPU pipeline vs the CPU oracle on edge cases: the Poisson rates of the
reference's numerical-stability test (tests/test_bmnd_robustness.py:49-67:
rates 0, 0.05, 5 and 1e6 on an 8x8 image, a single patch position), an
all-zero volume, extreme count rates in a volume, search boxes clipped by a
thin axis and an odd-sized image.  The bar is the same as
tests/test_gpu_parity.py: identical match lists in both stages and a
bit-identical output."""

import numpy as np
import pytest

import oracle
from bmnd_amd import BM3DProfile, BM4DProfile, NoiseModel, bmnd, phantoms

from test_gpu_parity import _compare

pytestmark = pytest.mark.gpu


def _assert_identical(stats):
    assert stats["ht_groups_differing"] == 0
    assert stats["wiener_groups_differing"] == 0
    assert stats["pilot_bits_differing"] == 0
    assert stats["out_bits_differing"] == 0


def _api_matches_oracle(noisy, prof, ndim):
    """Through the public bmnd() (which, like the reference, returns zeros
    for an all-zero input without running the stages)."""
    out = bmnd(noisy, prof)
    o_out = oracle.bmnd_poisson(noisy, oracle.default_params(ndim, noisy.shape))[0]
    assert out.shape == noisy.shape and out.dtype == np.float32
    assert np.all(np.isfinite(out))
    assert np.array_equal(out.view(np.int32), o_out.view(np.int32))


@pytest.mark.parametrize("rate", [0.0, 0.05, 5.0, 1.0e6])
def test_reference_stability_rates_on_one_patch_image(rate):
    noisy = np.random.default_rng(0).poisson(rate, size=(8, 8)).astype(np.float32)
    prof = BM3DProfile(noise_model=NoiseModel.POISSON)
    if not np.any(noisy):
        _api_matches_oracle(noisy, prof, 2)
        return
    stats, out, _ = _compare(noisy, prof, oracle.default_params(2, noisy.shape), "auto-fast")
    _assert_identical(stats)
    assert np.all(np.isfinite(out))


def test_all_zero_volume():
    _api_matches_oracle(np.zeros((20, 24, 28), dtype=np.float32),
                        BM4DProfile(noise_model=NoiseModel.POISSON), 3)


@pytest.mark.parametrize("peak", [0.05, 1.0e6])
def test_extreme_count_rates(peak):
    """Nearly all-zero counts, and counts far above the integer-table range
    (stage 1 then runs the fast certified matcher as well)."""
    _, noisy = phantoms.ellipsoid_volume((24, 26, 22), seed=11, peak=peak, n_ellipsoids=5,
                                         axes_range=(3.0, 10.0))
    prof = BM4DProfile(noise_model=NoiseModel.POISSON)
    stats, out, _ = _compare(noisy, prof, oracle.default_params(3, noisy.shape), "auto-fast")
    _assert_identical(stats)
    assert np.all(np.isfinite(out))


def test_thin_axis_clips_every_search_box():
    """Axis 0 shorter than a search box: every reference's box is clipped
    and the patch grid along it has two or three positions."""
    _, noisy = phantoms.ellipsoid_volume((7, 36, 33), seed=13, peak=10.0, n_ellipsoids=6,
                                         axes_range=(2.0, 12.0))
    prof = BM4DProfile(noise_model=NoiseModel.POISSON)
    stats, out, _ = _compare(noisy, prof, oracle.default_params(3, noisy.shape), "auto-fast")
    _assert_identical(stats)


def test_odd_sized_image():
    _, noisy = phantoms.blob_image(53, seed=17, peak=10.0)
    noisy = np.ascontiguousarray(noisy[:37, :53])
    prof = BM3DProfile(noise_model=NoiseModel.POISSON)
    stats, out, _ = _compare(noisy, prof, oracle.default_params(2, noisy.shape), "auto-fast")
    _assert_identical(stats)
