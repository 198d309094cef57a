"""The member-pair form of the exact overlap-aware planes, checked on the CPU.

The device kernel bmnd_hip_exact_std_pairs evaluates
  var[q,s] = sum_{i <= j, overlapping} c_q(i,j) K_ij[s],
  c_q(i,j) = G[q,i] G[q,j] (x2 for i < j),
  K_ij[s]  = sum_{p in overlap} var(x_i + p) S[s,p] S[s,p - (x_j - x_i)],
with each K_ij as three separable passes over the overlap region (tables
H_a^e[t,p] = A_a[t,p] A_a[t,p-e]).  Here that evaluation is restated in
float64 with the kernel's index arithmetic (pair list, overlap box, pass
layouts, the multiply-shift division) and held to the dense T Sigma T^T
oracle of the reference's variance tests (tests/test_variance_unit.py:46-86),
on overlapping and scattered members, 1-D/2-D/3-D blocks, and more planes
than one round of four."""

import numpy as np
import pytest

from bmnd_amd import transforms

from test_oracle_variance import _dense_exact


def _pairs_planes(vv, pos, block, g, spatial, planes):
    B = [1] * (3 - len(block)) + list(block)
    A = [np.ones((1, 1))] * (3 - len(block)) + [np.asarray(m, np.float64) for m in spatial]
    vv3 = vv.reshape((1,) * (3 - vv.ndim) + vv.shape)
    pos3 = np.concatenate([np.zeros((pos.shape[0], 3 - pos.shape[1]), np.int64), pos], 1)
    n = pos.shape[0]
    P = B[0] * B[1] * B[2]
    H = []
    for a in range(3):
        b = B[a]
        h = np.zeros((2 * b - 1, b, b))
        for e in range(2 * b - 1):
            for t in range(b):
                for p in range(b):
                    q = p - (e - (b - 1))
                    if 0 <= q < b:
                        h[e, t, p] = A[a][t, p] * A[a][t, q]
        H.append(h)
    vpatch = np.stack([vv3[o[0]:o[0] + B[0], o[1]:o[1] + B[1], o[2]:o[2] + B[2]].reshape(P)
                       for o in pos3])
    pairs = []
    for i in range(n):
        for j in range(i, n):
            d = pos3[j] - pos3[i]
            if all(abs(d[a]) < B[a] for a in range(3)):
                pairs.append((i, j, [int(d[a]) + B[a] - 1 for a in range(3)]))
    out = np.zeros((planes, P))
    for i, j, e in pairs:
        dd = [e[a] - (B[a] - 1) for a in range(3)]
        lo = [max(dd[a], 0) for a in range(3)]
        m = [B[a] - abs(dd[a]) for a in range(3)]
        mg1 = 65536 // m[1] + 1
        vi = vpatch[i]
        y2 = np.zeros(m[0] * m[1] * B[2])
        for t in range(m[0] * m[1] * B[2]):
            r, t2 = t // B[2], t % B[2]
            u0 = (r * mg1) >> 16
            u1 = r - u0 * m[1]
            row = vi[((lo[0] + u0) * B[1] + lo[1] + u1) * B[2]:][:B[2]]
            y2[t] = np.dot(H[2][e[2], t2], row)
        y1 = np.zeros(m[0] * B[1] * B[2])
        for t in range(m[0] * B[1] * B[2]):
            u0, t12 = t // (B[1] * B[2]), t % (B[1] * B[2])
            t1, t2 = t12 // B[2], t12 % B[2]
            y1[t] = sum(H[1][e[1], t1, lo[1] + u1] * y2[(u0 * m[1] + u1) * B[2] + t2]
                        for u1 in range(m[1]))
        K = np.zeros(P)
        for s in range(P):
            t0, t12 = s // (B[1] * B[2]), s % (B[1] * B[2])
            K[s] = sum(H[0][e[0], t0, lo[0] + u0] * y1[u0 * B[1] * B[2] + t12]
                       for u0 in range(m[0]))
        c = np.asarray(g, np.float64)[:planes, i] * np.asarray(g, np.float64)[:planes, j]
        out += np.outer(c * (1.0 if i == j else 2.0), K)
    return out.reshape((planes,) + tuple(block))


@pytest.mark.parametrize("case", ["3d-overlap", "3d-scattered", "2d-bior", "1d", "3d-k8"])
def test_pair_form_equals_dense_oracle(case):
    rng = np.random.default_rng({"3d-overlap": 0, "3d-scattered": 1, "2d-bior": 2, "1d": 3,
                                 "3d-k8": 4}[case])
    if case == "2d-bior":
        block, n = (4, 4), 8
        vv = rng.uniform(0.1, 9, size=(12, 13)).astype(np.float32)
        pos = rng.integers(0, 8, size=(n, 2)).astype(np.int64)
        f, _ = transforms.bior15_matrices(4)
        spatial = [f, f]
    elif case == "1d":
        block, n = (2,), 3
        vv = np.array([1.0, 2.0, 5.0, 3.0], np.float32)
        pos = np.array([[0], [1], [2]], np.int64)
        spatial = [transforms.dct_matrix(2)]
    else:
        block = (3, 3, 3)
        n = 16 if case == "3d-k8" else 8
        vv = rng.uniform(0.1, 20, size=(11, 10, 12)).astype(np.float32)
        hi = 3 if case != "3d-scattered" else 8
        pos = np.unique(np.stack([rng.integers(0, hi, n * 2) for _ in range(3)], -1), axis=0)
        pos = pos[rng.permutation(len(pos))[:n]].astype(np.int64)
        n = pos.shape[0]
        spatial = [transforms.dct_matrix(3)] * 3
    g = transforms.group_matrices(n)[0]
    planes = min(8 if case == "3d-k8" else 4, n)
    got = _pairs_planes(vv, pos, block, g, spatial, planes)
    dense = _dense_exact(vv, pos, block, g, spatial)[:planes]
    assert np.allclose(got, dense, rtol=1e-10, atol=1e-9)


def test_multiply_shift_division():
    # the kernel's r / m == (r * (65536 / m + 1)) >> 16 for its ranges
    for m in range(1, 9):
        mg = 65536 // m + 1
        for r in range(0, 8 * 64):
            assert (r * mg) >> 16 == r // m
