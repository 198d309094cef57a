"""The oracle's exact covariance-aware group risk (wiener.py:638-844) against
the dense definition the reference's own tests hold it to
(tests/test_variance_unit.py:89-200, 1111-1372):
  noise = trace(W Finv Filt Sigma_T Filt^T Finv^T),
  bias  = trace(W Finv (Filt - I) diag(S) (Filt - I)^T Finv^T),
with Sigma_T = F Sigma F^T the transformed entry covariance (Sigma[e,e'] =
var_v when entries e, e' fall on voxel v), kept exact over the first k planes
and replaced by sigma^2 on the diagonal of the later ones, Filt = diag(gain),
W = diag(w^2) per member, F / Finv the Kronecker analysis / synthesis
matrices.  Float64 here; the oracle sums float32 contributions, so the bar is
1e-5 relative.  The no-overlap case reduces to the diagonal risk
(tests/test_variance_unit.py:871-918)."""

import numpy as np
import pytest

import oracle
from bmnd_amd import transforms


def _dense_risk(vv, pos, block, gf, gi, sf, si, window, gain, planes, sigma, spow):
    n = pos.shape[0]
    local = np.stack(np.meshgrid(*[np.arange(b) for b in block], indexing="ij"),
                     -1).reshape(-1, len(block))
    P = local.shape[0]
    S = np.array([[1.0]])
    Si = np.array([[1.0]])
    for f, i in zip(sf, si):
        S = np.kron(S, np.asarray(f, np.float64))
        Si = np.kron(Si, np.asarray(i, np.float64))
    F = np.kron(np.asarray(gf, np.float64), S)
    Finv = np.kron(np.asarray(gi, np.float64), Si)
    coords = (pos[:, None, :] + local[None]).reshape(n * P, len(block))
    keys = [tuple(c) for c in coords]
    cov = np.zeros((n * P, n * P))
    for a in range(n * P):
        for b in range(n * P):
            if keys[a] == keys[b]:
                cov[a, b] = vv[keys[a]]
    T = F @ cov @ F.T
    ke = min(planes, n) * P
    H = np.zeros_like(T)
    H[:ke, :ke] = T[:ke, :ke]
    rest = np.arange(ke, n * P)
    H[rest, rest] = (np.asarray(sigma, np.float64).reshape(-1) ** 2)[rest]
    G = np.diag(np.asarray(gain, np.float64).reshape(-1))
    W = np.diag(np.tile(np.asarray(window, np.float64).reshape(-1) ** 2, n))
    noise = np.trace(W @ Finv @ G @ H @ G.T @ Finv.T)
    bias = 0.0
    if spow is not None:
        B = G - np.eye(n * P)
        bias = np.trace(W @ Finv @ B @ np.diag(np.asarray(spow, np.float64).reshape(-1)) @ B.T
                        @ Finv.T)
    return noise + bias


def _case(kind, seed):
    rng = np.random.default_rng(seed)
    if kind == "1d":
        vv = np.array([1.0, 2.0, 5.0], np.float32)
        pos = np.array([[0], [1]], np.int64)
        block = (2,)
        f = transforms.dct_matrix(2)
        sf, si = [f], [np.ascontiguousarray(f.T)]
        window = np.array([1.0, 0.5], np.float32)
    elif kind == "2d":
        vv = rng.uniform(0.2, 9.0, size=(10, 11)).astype(np.float32)
        n = 4
        pos = np.unique(rng.integers(0, 6, size=(3 * n, 2)), axis=0)[:n].astype(np.int64)
        block = (4, 4)
        f, i = transforms.bior15_matrices(4)
        sf, si = [f, f], [i, i]
        window = np.outer(np.kaiser(4, 2.0), np.kaiser(4, 2.0)).astype(np.float32)
    else:
        vv = rng.uniform(0.2, 9.0, size=(8, 9, 8)).astype(np.float32)
        n = 8
        pos = np.unique(rng.integers(0, 4, size=(3 * n, 3)), axis=0)
        pos = pos[rng.permutation(len(pos))[:n]].astype(np.int64)
        block = (3, 3, 3)
        f = transforms.dct_matrix(3)
        sf, si = [f] * 3, [np.ascontiguousarray(f.T)] * 3
        k1 = np.kaiser(3, 2.0)
        window = np.einsum("a,b,c->abc", k1, k1, k1).astype(np.float32)
    n = pos.shape[0]
    gf, gi = transforms.group_matrices(n)
    P = int(np.prod(block))
    gain = rng.uniform(0.0, 1.0, size=(n, P)).astype(np.float32)
    sigma = rng.uniform(0.3, 2.0, size=(n, P)).astype(np.float32)
    spow = rng.uniform(0.0, 3.0, size=(n, P)).astype(np.float32)
    return vv, pos, block, gf, gi, sf, si, window, gain, sigma, spow


@pytest.mark.parametrize("kind", ["1d", "2d", "3d"])
@pytest.mark.parametrize("planes", [1, 2, 4, 16])
@pytest.mark.parametrize("bias", [False, True])
def test_exact_group_risk_equals_dense_definition(kind, planes, bias):
    vv, pos, block, gf, gi, sf, si, window, gain, sigma, spow = _case(kind, 3)
    sp = spow if bias else None
    got = oracle.exact_group_risk(vv, block, pos, gf, gi, sf, si, window, planes, gain, sigma,
                                  sp)
    want = _dense_risk(vv, pos, block, gf, gi, sf, si, window, gain, planes, sigma, sp)
    assert got == pytest.approx(want, rel=1e-5)


def test_exact_group_risk_without_overlap_is_the_diagonal_risk():
    # tests/test_variance_unit.py:871-918: identity transforms, no overlap
    vv = np.array([1.0, 2.0, 3.0, 4.0], np.float32)
    pos = np.array([[0], [2]], np.int64)
    eye2 = np.eye(2, dtype=np.float32)
    gain = np.array([[0.0, 0.3], [0.6, 0.9]], np.float32)
    sigma = np.sqrt(np.array([[1.0, 2.0], [3.0, 4.0]], np.float32))
    spow = np.array([[0.5, 0.25], [1.0, 2.0]], np.float32)
    got = oracle.exact_group_risk(vv, (2,), pos, eye2, eye2, [eye2], [eye2], np.ones(2), 2, gain,
                                  sigma, spow)
    diag = float(np.sum(gain.astype(np.float64) ** 2 * sigma.astype(np.float64) ** 2
                        + (1.0 - gain.astype(np.float64)) ** 2 * spow))
    assert got == pytest.approx(diag, abs=1e-6)


def test_exact_group_risk_moves_with_overlap():
    # tests/test_variance_unit.py:1815-...: with overlap the exact risk differs
    # from the diagonal proxy built from the same noise stds
    vv, pos, block, gf, gi, sf, si, window, gain, sigma, spow = _case("3d", 5)
    exact = oracle.exact_group_risk(vv, block, pos, gf, gi, sf, si, window, pos.shape[0], gain,
                                    sigma, None)
    proxy = oracle.exact_group_risk(vv, block, pos, gf, gi, sf, si, window, 0, gain, sigma, None)
    assert abs(exact - proxy) > 1e-3 * abs(proxy)
