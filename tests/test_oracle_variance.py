"""Coefficient noise std: the reference's known answers
(tests/test_variance_unit.py) and the dense T Sigma T^T oracle it uses."""

import numpy as np
import pytest

import oracle
from bmnd_amd import transforms


def _dense_exact(var_volume, abs_pos, block, group_mat, spatial):
    """diag(T Sigma T^T) with Sigma[e,e'] = var_v when entries e, e' hit the
    same voxel v (tests/test_variance_unit.py:46-86), in float64."""
    n = abs_pos.shape[0]
    local = np.stack(np.meshgrid(*[np.arange(b) for b in block], indexing="ij"), -1).reshape(-1, len(block))
    P = local.shape[0]
    S = np.array([[1.0]])
    for b, m in zip(block, spatial):
        S = np.kron(S, np.eye(b) if m is None else np.asarray(m, np.float64))
    T = np.kron(np.asarray(group_mat, np.float64), S)
    coords = (abs_pos[:, None, :] + local[None]).reshape(n * P, len(block))
    keys = [tuple(c) for c in coords]
    cov = np.zeros((n * P, n * P))
    for i in range(n * P):
        for j in range(n * P):
            if keys[i] == keys[j]:
                cov[i, j] = var_volume[keys[i]]
    return np.diag(T @ cov @ T.T).reshape((n,) + tuple(block))


def test_exact_matches_dense_covariance_diagonal():
    # tests/test_variance_unit.py:746-775 (k = 2 and the over-full k = 3)
    vv = np.array([1.0, 2.0, 5.0], np.float32)
    pos = np.array([[0], [1]], np.int64)
    h = ((1.0 / np.sqrt(2.0)) * np.array([[1.0, 1.0], [1.0, -1.0]])).astype(np.float32)
    dense = _dense_exact(vv, pos, (2,), h, [h])
    for k in (2, 3):
        sig = oracle.noise_std(vv, (2,), pos, h, [h], k)
        assert np.allclose(sig.astype(np.float64) ** 2, dense, atol=1e-6)


def test_exact_equals_approximate_without_overlap():
    # tests/test_variance_unit.py:719-743
    vv = np.array([1.0, 2.0, 3.0, 4.0], np.float32)
    pos = np.array([[0], [2]], np.int64)
    h = ((1.0 / np.sqrt(2.0)) * np.array([[1.0, 1.0], [1.0, -1.0]])).astype(np.float32)
    approx = oracle.noise_std(vv, (2,), pos, h, [h], 0)
    exact = oracle.noise_std(vv, (2,), pos, h, [h], 2)
    assert np.allclose(exact, approx, atol=1e-6)


def test_limit_planes_known_answer():
    # tests/test_variance_unit.py:496-512
    ex = np.sqrt(np.array([[1.5], [0.5]], np.float32))
    ap = np.ones((2, 1), np.float32)
    assert np.allclose(oracle.limit_planes(ex, ap, 0), ap)
    assert np.allclose(oracle.limit_planes(ex, ap, 1) ** 2, [[1.5], [0.5]])
    assert np.allclose(oracle.limit_planes(ex[:1], ap, 1), oracle.limit_planes(ex, ap, 1))
    assert np.allclose(oracle.limit_planes(ex, ap, 2), ex)


def test_compact_k1_equals_limited_full_exact():
    # tests/test_variance_unit.py:817-868
    vv = np.arange(1.0, 7.0, dtype=np.float32)
    pos = np.arange(4, dtype=np.int64).reshape(-1, 1)
    h = transforms.haar_matrix(4)
    d = transforms.dct_matrix(2)
    approx = oracle.noise_std(vv, (2,), pos, h, [d], 0)
    full = oracle.noise_std(vv, (2,), pos, h, [d], 4)
    k1 = oracle.noise_std(vv, (2,), pos, h, [d], 1)
    expected = oracle.limit_planes(full, approx, 1)
    assert np.allclose(k1, expected, atol=1e-6)


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_exact_planes_match_dense_oracle_with_overlap(seed):
    rng = np.random.default_rng(seed)
    vv = rng.uniform(0.1, 20, size=(9, 10, 11)).astype(np.float32)
    block = (3, 3, 3)
    n = 8
    pos = np.stack([rng.integers(0, 10 - 3, n), rng.integers(0, 11 - 3, n),
                    rng.integers(0, 12 - 3, n)], -1).astype(np.int64)
    pos[:, 0] = np.minimum(pos[:, 0], 6)
    g = transforms.haar_matrix(n)
    dct = transforms.dct_matrix(3)
    ex = oracle.exact_variance(vv, block, pos, g, [dct, dct, dct], n)
    dense = _dense_exact(vv, pos, block, g, [dct, dct, dct])
    assert np.allclose(ex, dense, rtol=2e-5, atol=1e-4)


def _serial_exact(vv, pos, block, g, spatial, planes):
    """Straight float32 restatement of variance.py:950-975 with the stable
    (member-order) segment order: the reference test's serial loop."""
    n = pos.shape[0]
    P = int(np.prod(block))
    local = np.stack(np.meshgrid(*[np.arange(b) for b in block], indexing="ij"), -1).reshape(-1, len(block))
    S = np.array([[1.0]], np.float32)
    for b, m in zip(block, spatial):
        S = np.kron(S, np.asarray(m, np.float32)).astype(np.float32)
    E = np.kron(np.asarray(g, np.float32)[:planes], S).astype(np.float32).T  # (nP, planes*P)
    mn = pos.min(0)
    coords = [tuple(pos[i] - mn + local[p]) for i in range(n) for p in range(P)]
    order = sorted(range(n * P), key=lambda e: (coords[e], e))
    acc = np.zeros(planes * P, np.float32)
    pos_ = 0
    while pos_ < len(order):
        cell = coords[order[pos_]]
        contrib = np.zeros(planes * P, np.float32)
        while pos_ < len(order) and coords[order[pos_]] == cell:
            contrib += E[order[pos_]]
            pos_ += 1
        v = vv[tuple(np.array(cell) + mn)]
        acc += (v * contrib) * contrib
    return acc


@pytest.mark.parametrize("seed", [0, 1])
def test_exact_planes_bitwise_equal_serial_reduction(seed):
    # tests/test_variance_unit.py:778-814 pins the accumulation to a serial loop
    rng = np.random.default_rng(seed)
    vv = rng.uniform(0.1, 9, size=(12, 12)).astype(np.float32)
    block = (4, 4)
    n = 8
    pos = rng.integers(0, 9, size=(n, 2)).astype(np.int64)
    g = transforms.haar_matrix(n)
    f, _ = transforms.bior15_matrices(4)
    ex = oracle.exact_variance(vv, block, pos, g, [f, f], 4)
    ser = _serial_exact(vv, pos, block, g, [f, f], 4)
    assert ex.reshape(-1).tobytes() == ser.tobytes()
