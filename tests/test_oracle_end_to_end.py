"""End-to-end invariants of the oracle pipeline (the reference pins its
Poisson output only by invariants: tests/test_bmnd_robustness.py,
tests/test_bmnd_smoke.py, tests/test_bmnd_pipeline_unit.py)."""

import numpy as np
import pytest

import oracle
from bmnd_amd import phantoms


def test_readme_example_denoises():
    truth, noisy = phantoms.square_image(0)
    out, pilot = oracle.bmnd_poisson(noisy)
    assert out.shape == noisy.shape and out.dtype == np.float32
    assert np.all(np.isfinite(out))
    assert phantoms.psnr(out, truth) > phantoms.psnr(noisy, truth) + 15.0
    assert phantoms.psnr(pilot, truth) > phantoms.psnr(noisy, truth) + 15.0


def test_deterministic_and_thread_count_invariant():
    _, noisy = phantoms.ellipsoid_volume((20, 22, 24), seed=1, peak=8.0, n_ellipsoids=5,
                                         axes_range=(3.0, 9.0))
    a, _ = oracle.bmnd_poisson(noisy, nthreads=1)
    b, _ = oracle.bmnd_poisson(noisy, nthreads=6)
    c, _ = oracle.bmnd_poisson(noisy, nthreads=6)
    assert a.tobytes() == b.tobytes() == c.tobytes()


@pytest.mark.parametrize("rate", [0.05, 5.0, 1e6])
def test_extreme_rates_stay_finite(rate):
    # tests/test_bmnd_robustness.py:54-67
    noisy = np.random.default_rng(3).poisson(rate, size=(24, 24)).astype(np.float32)
    out, _ = oracle.bmnd_poisson(noisy)
    assert np.all(np.isfinite(out)) and out.shape == noisy.shape


def test_all_zero_volume_is_returned():
    z = np.zeros((16, 16), np.float32)
    out, pilot = oracle.bmnd_poisson(z)
    assert np.array_equal(out, z) and np.array_equal(pilot, z)


def test_match_lists_start_with_the_reference():
    _, noisy = phantoms.square_image(2)
    out, pilot, (hs, hc), (ws, wc) = oracle.bmnd_poisson(noisy, capture_matches=True)
    refs = oracle.schedule((57, 57), (8, 8), (3, 3), 2.0, 2)
    lin = refs[:, 0] * 57 + refs[:, 1]
    assert np.array_equal(hs[:, 0], lin) and np.array_equal(ws[:, 0], lin)
    for cnt in (hc, wc):
        assert np.all((cnt & (cnt - 1)) == 0) and np.all(cnt >= 2)
