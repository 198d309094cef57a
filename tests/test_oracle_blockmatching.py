"""Oracle block matching pinned by the reference's known answers
(tests/test_blockmatching_unit.py, tests/test_blockmatching_internal_unit.py)
and by an independent brute-force selection."""

import math

import numpy as np
import pytest

import oracle
from oracle import DIST_ANSCOMBE, DIST_DEVIANCE, DIST_PEARSON, DIST_SSD


def _legacy(patches, refs, counts, radius, max_m, thr, code=DIST_SSD, scale=1.0):
    """The reference's _batched_distance_blockmatch: min 1, no power of two,
    no tie canonicalisation (blockmatching.py:700-732)."""
    res = oracle.blockmatch(refs, counts, radius, max_m, 1, thr, patches_flat=patches,
                            pow2=False, stabilize=False, code=code, count_scale=scale)
    return res.lists()


def _group(patches, ref, counts, radius, max_m, min_m, thr, pow2=False):
    """The reference's blockmatch_group defaults: SSD, no stabilisation."""
    return oracle.blockmatch([ref], counts, radius, max_m, min_m, thr, patches_flat=patches,
                             pow2=pow2, stabilize=False, code=DIST_SSD)


def test_self_only_with_zero_radius():
    p = np.array([[0.0, 0.0], [1.0, 1.0], [2.0, 2.0]], np.float32)
    assert _legacy(p, [(0,), (1,), (2,)], (3,), (0,), 2, 1.0) == [[0], [1], [2]]


def test_includes_self_in_neighbours():
    p = np.array([[0.0, 0.0], [1.0, 1.0], [2.0, 2.0]], np.float32)
    assert _legacy(p, [(1,)], (3,), (1,), 2, 100.0) == [[1, 0]]


@pytest.mark.parametrize("code", [DIST_DEVIANCE, DIST_PEARSON, DIST_ANSCOMBE])
def test_poisson_distances_account_for_count_dependent_variance(code):
    # tests/test_blockmatching_unit.py:101-133
    p = np.array([[0.0, 2.0], [2.0, 2.0], [1.0, 1.0]], np.float32)
    assert _legacy(p, [(2,)], (3,), (2,), 3, 100.0, DIST_SSD) == [[2, 0, 1]]
    assert _legacy(p, [(2,)], (3,), (2,), 3, 100.0, code) == [[2, 1, 0]]


@pytest.mark.parametrize("code", [DIST_DEVIANCE, DIST_PEARSON, DIST_ANSCOMBE])
def test_poisson_distances_invariant_to_representation_scale(code):
    # tests/test_blockmatching_unit.py:141-180
    p = np.array([[11.0, 9.0], [10.0, 10.0], [20.0, 20.0]], np.float32)
    a = _legacy(p, [(1,)], (3,), (1,), 3, 0.2, code, 1.0)
    b = _legacy(7.0 * p, [(1,)], (3,), (1,), 3, 0.2, code, 7.0)
    assert a == [[1, 0]]
    assert b == a


def test_fills_to_min_stack_when_threshold_is_tight():
    # tests/test_blockmatching_internal_unit.py:96-113
    p = np.array([[0.0], [10.0], [20.0], [30.0]], np.float32)
    r = _group(p, (1,), (4,), (3,), 4, 3, 1e-8)
    assert int(r.accepted[0]) == 1
    assert r.lists() == [[1, 0, 2]]


def test_breaks_distance_ties_by_linear_index():
    # tests/test_blockmatching_internal_unit.py:135-151
    p = np.array([[1.0], [3.0], [2.0], [1.0], [3.0]], np.float32)
    assert _group(p, (2,), (5,), (2,), 5, 1, 10.0).lists() == [[2, 0, 1, 3, 4]]


def test_keeps_reference_first_when_nothing_is_accepted():
    # tests/test_blockmatching_internal_unit.py:154-171
    p = np.array([[0.0], [10.0], [20.0], [30.0]], np.float32)
    r = _group(p, (1,), (4,), (3,), 2, 1, -1.0)
    assert int(r.accepted[0]) == 0
    assert r.lists() == [[1]]


def test_power_of_two_after_reference_insertion():
    # tests/test_blockmatching_internal_unit.py:268-285
    p = np.array([[1.0], [3.0], [2.0], [1.0], [3.0]], np.float32)
    r = _group(p, (2,), (5,), (2,), 5, 2, 10.0, pow2=True)
    assert r.lists() == [[2, 0, 1, 3]]


def test_batch_and_thread_invariance():
    # tests/test_blockmatching_internal_unit.py:288-347 (the oracle has no
    # batches; its thread count is the analogue)
    rng = np.random.default_rng(0)
    vol = rng.poisson(5.0, size=(20, 21)).astype(np.float32)
    refs = np.stack(np.meshgrid(np.arange(0, 13, 3), np.arange(0, 14, 2), indexing="ij"), -1).reshape(-1, 2)
    kw = dict(volume=vol, block=(8, 8))
    a = oracle.blockmatch(refs, (13, 14), (5, 5), 16, 2, 2.9527, nthreads=1, **kw)
    b = oracle.blockmatch(refs, (13, 14), (5, 5), 16, 2, 2.9527, nthreads=7, **kw)
    assert a.lists() == b.lists()
    assert np.array_equal(a.accepted, b.accepted)


def test_canonical_key_grid():
    # blockmatching.py:957-964: q = 32*eps32*2^floor(log2(max(|d|,1)))
    q0 = 32 * 2.0 ** -23
    assert oracle.canonical_key(0.0) == 0.0
    assert oracle.canonical_key(q0 * 0.49) == 0.0
    assert oracle.canonical_key(q0 * 0.51) == pytest.approx(q0, rel=0, abs=0)
    # half to even: 0.5 q -> 0, 1.5 q -> 2 q
    assert oracle.canonical_key(np.float32(q0 * 0.5)) == 0.0
    assert oracle.canonical_key(np.float32(q0 * 1.5)) == np.float32(2 * q0)
    d = np.float32(3.0)
    q1 = 32 * 2.0 ** -23 * 2.0
    assert oracle.canonical_key(d) == np.float32(round(3.0 / q1) * q1)


def test_canonical_key_float32_scaling_form():
    """The matchers' device key computes the tie grid in float32 as two
    power-of-two scalings, ldexp(rint(ldexp(d, 19 - e)), e - 19) with e the
    frexp exponent of max(|d|, 1); it must give the bits of the reference's
    float64 form float32(rint(d / q) * q) (blockmatching.py:957-964) for
    every float32 d.  Checked on 2^24 random bit patterns, every exponent's
    grid boundaries and the halfway points, whose ties round to even."""
    rng = np.random.default_rng(5)
    bits = [rng.integers(0, 1 << 32, 1 << 24, dtype=np.uint64).astype(np.uint32)]
    for ex in range(0, 255):
        base = np.uint32(ex << 23)
        bits.append(base + np.arange(0, 64, dtype=np.uint32))
        bits.append(base + np.uint32((1 << 23) - 64) + np.arange(64, dtype=np.uint32))
    d = np.concatenate(bits).view(np.float32)
    d = np.concatenate([d, -d])
    q0 = np.float32(32 * 2.0 ** -23)
    halves = (np.arange(1, 4096, dtype=np.float32) + np.float32(0.5)) * q0
    d = np.concatenate([d, halves, halves * np.float32(2.0 ** 40)])
    d = d[np.isfinite(d)]
    mag = np.maximum(np.abs(d.astype(np.float64)), 1.0)
    e = np.frexp(mag)[1]
    q = np.ldexp(2.0 ** -18, e - 1)
    with np.errstate(over="ignore"):
        want = (np.rint(d.astype(np.float64) / q) * q).astype(np.float32)
        e32 = np.frexp(np.maximum(np.abs(d), np.float32(1.0)))[1]
        got = np.ldexp(np.rint(np.ldexp(d, 19 - e32)), e32 - 19).astype(np.float32)
    assert got.dtype == np.float32
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def _deviance_f64(r, c):
    """Term-by-term restatement of blockmatching.py:338-371 in Python."""
    d = 0.0
    for rr, cc in zip(r.astype(np.float64), c.astype(np.float64)):
        rr, cc = max(rr, 0.0), max(cc, 0.0)
        s = rr + cc
        if s > 0:
            t = 0.0
            if rr > 0:
                t += rr * math.log(2.0 * rr / s)
            if cc > 0:
                t += cc * math.log(2.0 * cc / s)
            d += 2.0 * t
    d = max(d, 0.0)
    return np.float32(d / r.size)


def test_patch_distance_matches_python_restatement():
    rng = np.random.default_rng(1)
    for _ in range(200):
        r = rng.poisson(rng.uniform(0.1, 30), size=64).astype(np.float32)
        c = rng.poisson(rng.uniform(0.1, 30), size=64).astype(np.float32)
        assert oracle.patch_distance(r, c) == _deviance_f64(r, c)
    z = np.zeros(8, np.float32)
    assert oracle.patch_distance(z, z) == 0.0
    x = rng.uniform(0, 9, 8).astype(np.float32)
    assert oracle.patch_distance(x, x) == 0.0


def _brute_force_select(vol, block, ref, radius, max_m, min_m, thr):
    """Independent selection: all candidates, sort by (key, lin), split by
    threshold key, fill, reference first, power-of-two."""
    counts = tuple(vol.shape[d] - block[d] + 1 for d in range(vol.ndim))
    lo = [max(ref[d] - radius[d], 0) for d in range(vol.ndim)]
    hi = [min(ref[d] + radius[d], counts[d] - 1) for d in range(vol.ndim)]
    rp = vol[tuple(slice(ref[d], ref[d] + block[d]) for d in range(vol.ndim))].ravel()
    cands = []
    for pos in np.ndindex(*[hi[d] - lo[d] + 1 for d in range(vol.ndim)]):
        o = [lo[d] + pos[d] for d in range(vol.ndim)]
        cp = vol[tuple(slice(o[d], o[d] + block[d]) for d in range(vol.ndim))].ravel()
        key = oracle.canonical_key(_deviance_f64(rp, cp))
        cands.append((key, int(np.ravel_multi_index(o, counts))))
    tkey = oracle.canonical_key(np.float32(thr))
    acc = sorted(c for c in cands if c[0] < tkey)
    rej = sorted(c for c in cands if not c[0] < tkey)
    sel = [c[1] for c in acc[:max_m]]
    need = min(min_m, max_m)
    if len(sel) < need:
        sel += [c[1] for c in rej[: need - len(sel)]]
    center = int(np.ravel_multi_index(ref, counts))
    out = [center] + [s for s in sel if s != center]
    out = out[:max_m]
    t = 1 << (len(out).bit_length() - 1)
    t = max(t, min_m)
    return out[: min(t, len(out))], len(acc)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_fused_matcher_equals_brute_force(seed):
    rng = np.random.default_rng(seed)
    lam = rng.uniform(0.5, 12.0, size=(14, 15))
    vol = rng.poisson(lam).astype(np.float32)
    block, radius, max_m, min_m, thr = (4, 4), (4, 5), 8, 2, 0.9
    counts = (11, 12)
    refs = [(0, 0), (5, 6), (10, 11), (3, 9), (7, 2)]
    got = oracle.blockmatch(refs, counts, radius, max_m, min_m, thr, volume=vol, block=block)
    for i, ref in enumerate(refs):
        exp, nacc = _brute_force_select(vol, block, ref, radius, max_m, min_m, thr)
        assert got.lists()[i] == exp
        assert int(got.accepted[i]) == nacc
