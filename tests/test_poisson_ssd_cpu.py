"""The SSD distance under the Poisson model: stage-1 matching with the Poisson
gamma correction (blockmatching.py:432-443, 1140-1145; bmndalgo.py:292-294,
321-322), pinned by the reference's known answer
(tests/test_blockmatching_internal_unit.py:203-229) and by an independent
selection that sorts by a total order instead of inserting candidates one by
one."""

import numpy as np
import pytest

import oracle


def test_reference_known_answer():
    # patches [0] and [2], reference 0, threshold 1, gamma 2: the candidate's
    # SSD 4 less 2 * (0 + 2) is 0, accepted; uncorrected it is rejected
    vol = np.array([0.0, 2.0], np.float32)
    plain = oracle.blockmatch_ssd_poisson(vol, (1,), (1,), [(0,)], 1.0, 2, 1, 0.0, pow2=False)
    corrected = oracle.blockmatch_ssd_poisson(vol, (1,), (1,), [(0,)], 1.0, 2, 1, 2.0, pow2=False)
    assert plain[0][0, :plain[1][0]].tolist() == [0]
    assert corrected[0][0, :corrected[1][0]].tolist() == [0, 1]
    assert int(corrected[2][0]) == 2


def _total_order_select(vol, block, ref, radius, max_m, min_m, thr, gamma, scale, pow2=True):
    """Selection by a total order.  The corrected distance K is a float64 and
    the buffers keep S = float32(K); a later candidate goes before an earlier
    one when K_later < S_earlier.  For candidates of one S value that puts
    those with K < S first, the later ones first, then those with K >= S in
    index order; across S values it is the order of S.  Sorting by
    (S, K >= S, +-index) and keeping the first ones gives the insertion's
    result."""
    counts = tuple(vol.shape[d] - block[d] + 1 for d in range(vol.ndim))
    lo = [max(ref[d] - radius[d], 0) for d in range(vol.ndim)]
    hi = [min(ref[d] + radius[d], counts[d] - 1) for d in range(vol.ndim)]
    rp = vol[tuple(slice(ref[d], ref[d] + block[d]) for d in range(vol.ndim))].ravel()
    center = int(np.ravel_multi_index(ref, counts))
    g32 = float(np.float32(gamma))
    t64 = float(np.float32(thr))
    acc, rej = [], []
    for pos in np.ndindex(*[hi[d] - lo[d] + 1 for d in range(vol.ndim)]):
        o = [lo[d] + pos[d] for d in range(vol.ndim)]
        cp = vol[tuple(slice(o[d], o[d] + block[d]) for d in range(vol.ndim))].ravel()
        lin = int(np.ravel_multi_index(o, counts))
        d32 = np.float32(0.0)
        bias = 0.0
        for r, c in zip(rp, cp):
            diff = np.float32(c - r)
            d32 = np.float32(d32 + np.float32(diff * diff))
            pooled = np.float32(r + c)
            if pooled > 0:
                bias += scale * float(pooled)
        k = float(d32)
        if lin != center:
            k = k - g32 * bias
        s = float(np.float32(k))
        g = k >= s
        entry = (s, int(g), lin if g else -lin, lin)
        (acc if k < t64 else rej).append(entry)
    acc.sort()
    rej.sort()
    sel = [e[3] for e in acc[:max_m]]
    need = min(min_m, max_m)
    if len(sel) < need:
        sel += [e[3] for e in rej[: need - len(sel)]]
    out = [center] + [s for s in sel if s != center]
    out = out[:max_m]
    if not pow2:
        return out, len(acc)
    t = 1 << (len(out).bit_length() - 1)
    t = max(t, min_m)
    return out[: min(t, len(out))], len(acc)


@pytest.mark.parametrize("seed,scale,gamma", [(0, 1.0, 3.0), (1, 1.0, 0.5), (2, 0.37, 3.0),
                                              (3, 2.5, 1.25)])
def test_matches_total_order_selection(seed, scale, gamma):
    rng = np.random.default_rng(seed)
    lam = rng.uniform(0.5, 12.0, size=(13, 14))
    vol = (rng.poisson(lam) * scale).astype(np.float32)
    if seed % 2:
        vol = vol + rng.uniform(-0.3, 0.3, vol.shape).astype(np.float32)
    block, radius, max_m, min_m = (4, 4), (4, 5), 8, 2
    thr = float(np.float32(2.9527 * float(vol.max() - vol.min()) ** 2 * 0.02))
    counts = (10, 11)
    refs = [(0, 0), (5, 6), (9, 10), (3, 9), (7, 2)]
    sel, cnt, acc = oracle.blockmatch_ssd_poisson(vol, block, radius, refs, thr, max_m, min_m,
                                                  gamma, count_scale=scale)
    for i, ref in enumerate(refs):
        exp, nacc = _total_order_select(vol, block, ref, radius, max_m, min_m, thr, gamma, scale)
        assert sel[i, :cnt[i]].tolist() == exp
        assert int(acc[i]) == nacc


def _mixed_pair(rng, scale):
    """Two one-voxel candidates c1 (earlier) and c2 (later) against a zero
    reference whose corrected distances K = f32(c^2) - scale * c (gamma 1)
    round to one float32 S with K2 < S <= K1: the reference's insertion then
    puts the later candidate first.  Found among random values by grouping
    them by S."""
    c = rng.uniform(2.0, 3.0, 400000).astype(np.float32)
    k = (c * c).astype(np.float64) - scale * c.astype(np.float64)
    s = k.astype(np.float32)
    order = np.argsort(s, kind="stable")
    s_sorted, k_sorted = s[order], k[order]
    for i in range(len(order) - 1):
        if s_sorted[i] != s_sorted[i + 1]:
            continue
        a, b = order[i], order[i + 1]
        lo, hi = (a, b) if k[a] < k[b] else (b, a)
        if k[lo] < float(s[lo]) <= k[hi]:
            return c[hi], c[lo]
    raise AssertionError("no pair found")


def test_mixed_precision_order():
    scale = 0.1
    c1, c2 = _mixed_pair(np.random.default_rng(5), scale)
    vol = np.array([0.0, c1, c2], np.float32)
    thr = 100.0
    sel, cnt, _ = oracle.blockmatch_ssd_poisson(vol, (1,), (2,), [(0,)], thr, 4, 1, 1.0,
                                                count_scale=scale, pow2=False)
    # the float32-rounded distances are equal; in index order c1 would come
    # first, the float64 comparison against c1's stored float32 puts c2 first
    assert sel[0, :cnt[0]].tolist() == [0, 2, 1]
    exp, _ = _total_order_select(vol, (1,), (0,), (2,), 4, 1, thr, 1.0, scale, pow2=False)
    assert exp == [0, 2, 1]


def test_zero_gamma_is_plain_ssd():
    rng = np.random.default_rng(7)
    vol = rng.poisson(4.0, size=(12, 12)).astype(np.float32)
    refs = [(0, 0), (4, 4), (8, 8)]
    a = oracle.blockmatch_ssd_poisson(vol, (4, 4), (3, 3), refs, 30.0, 8, 1, 0.0)
    b = oracle.blockmatch_ssd(vol, (4, 4), (3, 3), refs, 30.0, 8, 1)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
