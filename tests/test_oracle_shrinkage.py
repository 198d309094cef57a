"""Thresholding, Wiener gain, weights and scatter against the reference's
closed forms and known answers."""

import numpy as np
import pytest

import oracle
from bmnd_amd import transforms


def test_wiener_gain_variance_scaled_closed_form():
    # tests/compat/test_component_semantics.py:54-77 (scale 1.0)
    ref = np.array([0.0, 0.1, 0.4, 1.5], np.float32)
    sig = np.full(4, 0.2, np.float32)
    coef = np.ones(4, np.float32)
    out, gain = oracle.wiener_shrink(coef, ref, sig, mu=1.0)
    signal = np.maximum(ref ** 2 - sig ** 2, 0.0)
    np.testing.assert_allclose(gain, signal / (signal + sig ** 2), rtol=1e-6, atol=1e-7)
    np.testing.assert_array_equal(out, gain)


def test_wiener_gain_scaled_noise():
    ref = np.array([0.0, 0.3, 0.6, 1.5], np.float32)
    sig = np.full(4, 0.2, np.float32)
    _, gain = oracle.wiener_shrink(np.ones(4, np.float32), ref, sig, mu=2.0)
    signal = np.maximum(ref ** 2 - 2.0 * sig ** 2, 0.0)
    np.testing.assert_allclose(gain, signal / (signal + 2.0 * sig ** 2), rtol=1e-6, atol=1e-7)
    assert np.all((gain >= 0) & (gain <= 1))


def test_hard_threshold_tolerance():
    # poisson.py:30-46: keep iff |C| > 3 sigma + 2^-16 max(|C|, 3 sigma, 1)
    sig = np.full(6, 1.0, np.float32)
    thr = 3.0
    tol = 2.0 ** -16 * 3.0
    c = np.array([3.0, 3.0 + tol * 0.5, 3.0 + tol * 2, -4.0, 0.0, -3.0 - tol * 4], np.float32)
    out, att = oracle.ht_shrink(c, sig, thr)
    assert att.tolist() == [0.0, 0.0, 1.0, 1.0, 0.0, 1.0]
    assert np.array_equal(out, c * att)


def test_invert_aggregation_risks_known_answer():
    # tests/test_aggregationweights_unit.py:315-324: risk 0 -> 1e10, risk 2 -> 0.5
    g = np.eye(2, dtype=np.float32)
    att = np.array([[0.0], [1.0]], np.float32)
    sig = np.array([[1.0], [np.sqrt(2.0)]], np.float32)
    w = oracle.patch_weights(att, sig, g)
    assert w[0] == np.float32(1e10)
    assert w[1] == pytest.approx(0.5, rel=1e-6)


def test_patch_risks_match_dense_metric():
    # tests/test_aggregationweights_unit.py:73-118 (coefficient domain)
    rng = np.random.default_rng(0)
    n, p = 4, 6
    ginv = transforms.haar_matrix(n).T.copy()
    att = rng.uniform(0, 1, (n, p)).astype(np.float32)
    sig = rng.uniform(0.1, 2, (n, p)).astype(np.float32)
    w = oracle.patch_weights(att, sig, ginv)
    err = att.astype(np.float64) ** 2 * (sig * sig).astype(np.float32).astype(np.float64)
    risk = (ginv.astype(np.float64) ** 2) @ err.sum(1)
    np.testing.assert_allclose(w, (1.0 / (risk + 1e-10)).astype(np.float32), rtol=1e-6)


def test_scatter_weighted_overlap_known_answer():
    # tests/test_patchaggregation_internal_unit.py:81-108
    num = np.zeros((3, 3), np.float32)
    den = np.zeros((3, 3), np.float32)
    filt = np.array([[1.0, 2.0, 3.0, 4.0], [10.0, 20.0, 30.0, 40.0]], np.float32)
    oracle.scatter_add(num, den, (2, 2), np.ones(4, np.float32), filt,
                       np.array([1.0, 0.5], np.float32), np.array([0, 3], np.int64))
    np.testing.assert_array_equal(num, [[1, 2, 0], [3, 9, 10], [0, 15, 20]])
    np.testing.assert_array_equal(den, [[1, 1, 0], [1, 1.5, 0.5], [0, 0.5, 0.5]])


def test_scatter_uses_window_once():
    # tests/test_patchaggregation_internal_unit.py:111-127
    num = np.zeros((2, 2), np.float32)
    den = np.zeros((2, 2), np.float32)
    win = np.array([1.0, 0.5, 0.25, 0.75], np.float32)
    oracle.scatter_add(num, den, (2, 2), win, np.full((1, 4), 3.0, np.float32),
                       np.array([2.0], np.float32), np.array([0], np.int64))
    assert np.allclose(num.ravel(), 6.0 * win)
    assert np.allclose(den.ravel(), 2.0 * win)
