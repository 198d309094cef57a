"""The FP64 build of the oracle (the same algorithm with every float32
quantity in float64) against the float32 oracle, which the GPU matches bit
for bit: the error of the float32 path against a float64 run.  The stage-1
pilots agree to about 1e-6 relative; the outputs to well under 1e-3 relative
(stage-2 matching on the slightly different pilots may pick a different
neighbour for a few groups) and within 0.01 dB PSNR."""

import numpy as np

import oracle
from bmnd_amd import phantoms


def _compare(noisy, truth):
    o32, p32 = oracle.bmnd_poisson(noisy)
    o64, p64 = oracle.bmnd_poisson_f64(noisy)
    assert o64.dtype == np.float64 and o64.shape == noisy.shape
    rel_out = float(np.max(np.abs(o32 - o64)) / np.max(np.abs(o64)))
    rel_pilot = float(np.max(np.abs(p32 - p64)) / np.max(np.abs(p64)))
    gap = abs(phantoms.psnr(o32, truth) - phantoms.psnr(o64, truth))
    return rel_out, rel_pilot, gap


def test_fp64_run_of_the_readme_image():
    truth, noisy = phantoms.square_image(0)
    rel_out, rel_pilot, gap = _compare(noisy, truth)
    assert rel_pilot < 1e-5 and rel_out < 1e-3 and gap < 0.01


def test_fp64_run_of_a_small_volume():
    truth, noisy = phantoms.ellipsoid_volume((28, 32, 24), seed=3, peak=10.0, n_ellipsoids=6,
                                             axes_range=(3.0, 10.0))
    rel_out, rel_pilot, gap = _compare(noisy, truth)
    assert rel_pilot < 1e-5 and rel_out < 1e-3 and gap < 0.01


def test_fp64_matrices_are_float64_constructions():
    """The FP64 build's DCT is orthonormal to float64 precision (a float32
    table would leave ~1e-7 residuals)."""
    m = oracle._spatial_matrices_f64("dct", (8,))[:64].reshape(8, 8)
    assert np.max(np.abs(m @ m.T - np.eye(8))) < 1e-14
