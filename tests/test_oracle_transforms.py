"""Basis matrices and windows: oracle (C) vs bmnd_amd (numpy, built the
reference's way), plus the mathematical properties that fix them."""

import numpy as np
import pytest

import oracle
from bmnd_amd import transforms
from bmnd_amd.utils import kaiser_window_nd


@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 32])
def test_haar_bitwise_and_orthonormal(n):
    a = oracle.haar(n)
    b = transforms.haar_matrix(n)
    assert a.tobytes() == b.tobytes()  # signed zeros included
    assert np.allclose(a.astype(np.float64) @ a.T.astype(np.float64), np.eye(n), atol=1e-6)
    # hierarchical order: DC first, then coarsest detail (transforms.py:393-408)
    assert np.all(a[0] > 0)
    if n >= 2:
        assert np.all(a[1, : n // 2] > 0) and np.all(a[1, n // 2:] < 0)


def test_haar_size_four_known_answer():
    # tests/test_transforms_extended_unit.py:84-111
    s = np.float32(1 / np.sqrt(2))
    h = transforms.haar_matrix(4)
    assert np.allclose(h, [[s * s] * 4, [s * s, s * s, -s * s, -s * s], [s, -s, 0, 0], [0, 0, s, -s]])


@pytest.mark.parametrize("n", [2, 3, 4, 5, 7, 8, 16])
def test_dct_bitwise_and_orthonormal(n):
    f, i = oracle.dct(n)
    g = transforms.dct_matrix(n)
    assert f.tobytes() == g.tobytes()
    assert i.tobytes() == np.ascontiguousarray(g.T).tobytes()
    assert np.allclose(g.astype(np.float64) @ g.T.astype(np.float64), np.eye(n), atol=1e-6)


@pytest.mark.parametrize("n", [4, 8])
def test_bior15_inverse(n):
    f, i = oracle.bior15(n)
    g, gi = transforms.bior15_matrices(n)
    assert f.tobytes() == g.tobytes()
    assert np.allclose(f.astype(np.float64) @ i.astype(np.float64), np.eye(n), atol=1e-6)
    # the C LU and numpy's LAPACK agree except in entries that are exactly
    # zero in exact arithmetic, where both hold rounding noise of ~1e-17
    big = np.abs(gi) > 1e-12
    assert np.array_equal(i[big], gi[big])
    assert np.all(np.abs(i[~big]) < 1e-15)
    # the matrices the pipelines use come from numpy, as in the reference
    ref = oracle.reference_spatial_matrices("bior", (n,))
    assert ref[n * n:].tobytes() == gi.tobytes()


def test_bior15_four_is_a_haar_like_orthogonal_basis():
    f, _ = transforms.bior15_matrices(4)
    rows = f.astype(np.float64) @ f.T.astype(np.float64)
    assert np.allclose(rows, np.diag(np.diag(rows)), atol=1e-7)


@pytest.mark.parametrize("m", list(range(1, 17)))
def test_kaiser_matches_numpy(m):
    assert np.allclose(oracle.kaiser(m, 2.0), np.kaiser(m, 2.0), rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("block", [(4, 4, 4), (5, 5, 5), (8, 8), (3, 7), (1, 6, 2)])
def test_kaiser_window_bitwise(block):
    # tests/test_utils_unit.py:18-24 pins the window to np.kaiser
    assert oracle.kaiser_nd(block, 2.0).tobytes() == kaiser_window_nd(block, 2.0).tobytes()


def test_separable_transform_round_trip():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(8, 4, 4, 4)).astype(np.float32)
    g = transforms.haar_matrix(8)
    f, i = transforms.bior15_matrices(4)
    y = oracle.apply_transform(x, [g, f, f, f])
    z = oracle.apply_transform(y, [np.ascontiguousarray(g.T), i, i, i])
    assert np.allclose(z, x, atol=1e-5)


def test_separable_transform_equals_kron():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(4, 5, 5)).astype(np.float32)
    g = transforms.haar_matrix(4)
    d = transforms.dct_matrix(5)
    y = oracle.apply_transform(x, [g, d, d])
    full = np.kron(g.astype(np.float64), np.kron(d.astype(np.float64), d.astype(np.float64)))
    assert np.allclose(y.reshape(-1), full @ x.reshape(-1).astype(np.float64), atol=1e-5)
