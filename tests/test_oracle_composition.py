"""Pins of the oracle's stage composition (which volume feeds which
component), on the CPU.

`recompose` rebuilds both stages group by group from the reference's
orchestration (bmndalgo.py:213-633 for stage 1, 635-1046 for stage 2),
calling only oracle components that are pinned on their own (schedule,
matcher, transforms, exact variance, shrinkage, weights, scatter; see the
other test_oracle_*.py files).  Written from the reference's text, it fixes
the routing the C pipeline (oracle_bmnd_poisson) must follow: the variance
model of each stage, the volume each stage matches on, the coefficients the
Wiener gain comes from and the fallback of the normalisation.  The C
pipeline must equal it bit for bit.

The other tests restate the reference's pipeline-level checks
(tests/test_bmnd_pipeline_unit.py): stage-2 matching on the pilot (:913-936),
the k-saturation identity (:236-258), noise reduction with k = 0
(:206-233), plus a count-scale equivariance check and the measured effect of
the union-voxel entry order (numba's quicksort argsort, variance.py:804,
against the stable order the oracle and the device use).
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle


def _spatial(kind: str, block) -> tuple[list, list]:
    flat = oracle.reference_spatial_matrices(kind, block)
    fwd, inv, off = [], [], 0
    for b in block:
        fwd.append(flat[off:off + b * b].reshape(b, b))
        off += b * b
    for b in block:
        inv.append(flat[off:off + b * b].reshape(b, b))
        off += b * b
    return fwd, inv


def _variance_model(x: np.ndarray, scale: float, floor: float) -> np.ndarray:
    # variance.py:581-585, numpy's own expression
    nonnegative = np.maximum(np.asarray(x, dtype=np.float32), 0.0)
    return np.maximum(scale * nonnegative, floor).astype(np.float32, copy=False)


def _patches(x: np.ndarray, lins, counts, block) -> np.ndarray:
    out = []
    for lin in lins:
        o = np.unravel_index(int(lin), counts)
        out.append(x[tuple(slice(o[d], o[d] + block[d]) for d in range(x.ndim))])
    return np.stack(out).astype(np.float32)


def _stage(p, wiener: bool, match_vol, noisy, pilot, var_vol, fallback, gain_mode):
    """One stage, group by group, in schedule order."""
    nd = noisy.ndim
    block = tuple(p.wi_block[:nd] if wiener else p.ht_block[:nd])
    step = tuple(p.wi_step[:nd] if wiener else p.ht_step[:nd])
    radius = tuple(p.wi_radius[:nd] if wiener else p.ht_radius[:nd])
    counts = tuple(noisy.shape[d] - block[d] + 1 for d in range(nd))
    beta = p.wi_beta if wiener else p.ht_beta
    kaiser = oracle.kaiser_nd(block, beta)
    fwd_s, inv_s = _spatial("dct" if wiener else "bior", block)
    refs = oracle.schedule(counts, block, step, p.shift_density, p.schedule_density)
    thr = p.wi_threshold if wiener else p.ht_threshold
    mres = oracle.blockmatch(refs, counts, radius, p.wi_max if wiener else p.ht_max,
                             p.wi_min if wiener else p.ht_min, thr, volume=match_vol,
                             block=block, count_scale=p.count_scale)
    num = np.zeros(noisy.shape, np.float32)
    den = np.zeros(noisy.shape, np.float32)
    for g in range(len(refs)):
        n = int(mres.counts[g])
        if n == 0:
            continue
        lins = mres.selected[g, :n]
        apos = np.stack(np.unravel_index(lins, counts), axis=1).astype(np.int64)
        G = oracle.haar(n)
        Ginv = np.ascontiguousarray(G.T)  # transforms.py:430-431
        coef = oracle.apply_transform(_patches(noisy, lins, counts, block), [G] + fwd_s)
        sigma = oracle.noise_std(var_vol, block, apos, G, fwd_s, p.k)
        if wiener:
            # the gain from the pilot's coefficients (bmndalgo.py:771-785, 878-891)
            rcoef = oracle.apply_transform(_patches(pilot, lins, counts, block), [G] + fwd_s)
            coef, att = oracle.wiener_shrink(coef, rcoef, sigma, p.wi_mu, gain_mode)
        else:
            coef, att = oracle.ht_shrink(coef, sigma, p.ht_lambda, p.count_scale)
        filt = oracle.apply_transform(coef, [Ginv] + inv_s)
        w = oracle.patch_weights(att, sigma, Ginv)
        oracle.scatter_add(num, den, block, kaiser, filt, w, lins)
    est = np.where(den > 0, num / (den + 1e-10), fallback)  # bmndalgo.py:626-631
    return est.astype(np.float32), mres


def recompose(vol, p, *, match_obs=False, var_obs=False, gain_mode="variance-scaled"):
    """Both stages as the reference orchestrates them: returns (out, pilot,
    stage-1 matches, stage-2 matches)."""
    vol = np.ascontiguousarray(vol, dtype=np.float32)
    var1 = _variance_model(vol, p.count_scale, p.variance_floor)
    pilot, m1 = _stage(p, False, vol, vol, None, var1, vol, gain_mode)
    # bmndalgo.py:635-646: the Wiener variance model from the pilot unless the
    # observation is its source; :714-720: matching on the pilot unless the
    # observation is the match source
    var2 = var1 if var_obs else _variance_model(pilot, p.count_scale, p.variance_floor)
    out, m2 = _stage(p, True, vol if match_obs else pilot, vol, pilot, var2, vol, gain_mode)
    return out, pilot, m1, m2


def _image(seed=3, size=32):
    truth = np.full((size, size), 4.0, np.float32)
    truth[8:24, 6:20] = 15.0
    return np.random.default_rng(seed).poisson(truth).astype(np.float32)


def _volume(seed=5, side=12):
    truth = np.full((side,) * 3, 3.0, np.float32)
    truth[3:9, 2:8, 4:10] = 12.0
    return np.random.default_rng(seed).poisson(truth).astype(np.float32)


def _same_bits(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.int32),
                          np.asarray(b, np.float32).view(np.int32))


def _lists_equal(mres, sel, cnt):
    return (np.array_equal(mres.counts, cnt)
            and all(np.array_equal(mres.selected[g, :c], sel[g, :c]) for g, c in enumerate(cnt)))


@pytest.mark.parametrize("kind", ["image", "volume"])
def test_pipeline_equals_group_by_group_recomposition(kind):
    vol = _image() if kind == "image" else _volume()
    p = oracle.default_params(vol.ndim, vol.shape, nthreads=2)
    out, pilot, (s1, c1), (s2, c2) = oracle.bmnd_poisson(vol, p, capture_matches=True)
    r_out, r_pilot, m1, m2 = recompose(vol, p)
    assert _lists_equal(m1, s1, c1) and _lists_equal(m2, s2, c2)
    assert _same_bits(pilot, r_pilot)
    assert _same_bits(out, r_out)


@pytest.mark.parametrize("opts", [
    {"match_obs": True}, {"var_obs": True}, {"gain_mode": "classic"},
    {"gain_mode": "noise-floor"}, {"match_obs": True, "var_obs": True}])
def test_option_routing_equals_recomposition(opts):
    """The observation as the Wiener match / variance source and the gain
    modes route as the reference does (bmndalgo.py:635-646, 714-720, 878-885):
    the recomposition, which follows those lines, equals the C pipeline."""
    vol = _image(seed=11, size=30)
    p = oracle.default_params(vol.ndim, vol.shape, nthreads=2)
    p.wi_match_observation = int(opts.get("match_obs", False))
    p.wi_var_observation = int(opts.get("var_obs", False))
    p.wi_gain_mode = oracle.GAIN_MODES[opts.get("gain_mode", "variance-scaled")]
    out, pilot, _, (s2, c2) = oracle.bmnd_poisson(vol, p, capture_matches=True)
    r_out, r_pilot, _, m2 = recompose(vol, p, **opts)
    assert _lists_equal(m2, s2, c2)
    assert _same_bits(pilot, r_pilot) and _same_bits(out, r_out)


def test_wiener_stage_matches_on_the_pilot():
    """tests/test_bmnd_pipeline_unit.py:913-936 checks which patches the
    stage-2 matcher receives; here its lists equal the matcher run on the
    pilot at the Wiener geometry and differ from the lists on the
    observation (and the observation source gives those)."""
    vol = _image(seed=7)
    nd = vol.ndim
    p = oracle.default_params(nd, vol.shape, nthreads=2)
    _, pilot, _, (s2, c2) = oracle.bmnd_poisson(vol, p, capture_matches=True)
    block = tuple(p.wi_block[:nd])
    counts = tuple(vol.shape[d] - block[d] + 1 for d in range(nd))
    refs = oracle.schedule(counts, block, tuple(p.wi_step[:nd]))

    def match(x):
        return oracle.blockmatch(refs, counts, tuple(p.wi_radius[:nd]), p.wi_max, p.wi_min,
                                 p.wi_threshold, volume=x, block=block)
    on_pilot, on_obs = match(pilot), match(vol)
    assert _lists_equal(on_pilot, s2, c2)
    assert not _lists_equal(on_obs, s2, c2)
    p.wi_match_observation = 1
    _, _, _, (o2, oc2) = oracle.bmnd_poisson(vol, p, capture_matches=True)
    assert _lists_equal(on_obs, o2, oc2)


@pytest.mark.parametrize("kind", ["image", "volume"])
def test_variance_planes_saturate_at_the_stack_size(kind):
    """tests/test_bmnd_pipeline_unit.py:236-258: with stacks of at most 4
    members every plane is exact at k = 4, so k = 8 gives the same output
    bit for bit; k = 2 does not."""
    vol = _image(seed=29) if kind == "image" else _volume(seed=29)
    p = oracle.default_params(vol.ndim, vol.shape, nthreads=2)
    p.ht_max = p.wi_max = 4
    outs = {}
    for k in (2, 4, 8):
        p.k = k
        outs[k] = oracle.bmnd_poisson(vol, p)
    assert _same_bits(outs[4][0], outs[8][0]) and _same_bits(outs[4][1], outs[8][1])
    assert not _same_bits(outs[2][0], outs[4][0])


def test_k0_reduces_noise_on_a_constant_image():
    """tests/test_bmnd_pipeline_unit.py:206-233 (Poisson rows, default
    variance/coefficient/patch weights): an 8x8 Poisson(5) image with k = 0,
    4x4 blocks at step 4, radius 1 and stacks of at most 4 comes out closer
    to 5 than it went in.  (The reference's test profile uses an N-D DCT
    including the group axis; here the group axis keeps the Haar transform of
    the Poisson profiles, bior1.5 4x4 for HT and DCT 4x4 for Wiener.)"""
    vol = np.random.default_rng(71).poisson(5.0, size=(8, 8)).astype(np.float32)
    p = oracle.default_params(2, vol.shape, nthreads=1)
    for d in range(2):
        p.ht_block[d] = p.wi_block[d] = 4
        p.ht_step[d] = p.wi_step[d] = 4
        p.ht_radius[d] = p.wi_radius[d] = 1
    p.ht_max = p.wi_max = 4
    p.k = 0
    p._keep = (oracle.reference_spatial_matrices("bior", (4, 4)),
               oracle.reference_spatial_matrices("dct", (4, 4)))
    p.ht_spatial, p.wi_spatial = p._keep[0].ctypes.data, p._keep[1].ctypes.data
    out, _ = oracle.bmnd_poisson(vol, p)
    assert out.shape == vol.shape and np.isfinite(out).all()
    assert np.mean((out - 5.0) ** 2) < np.mean((vol - 5.0) ** 2)


@pytest.mark.parametrize("kind", ["image", "volume"])
def test_count_scale_equivariance(kind):
    """Data 2x with count scale 2 are the same counts: the matcher divides
    by the scale (blockmatching.py:338-371), the variance model multiplies by
    it (variance.py:581-585) and the hard threshold compares counts
    (poisson.py:30-46).  Stage-1 lists are identical, the estimates equal
    2x the unit-scale estimates up to the fixed 1e-10 terms (variance floor,
    weight and normalisation eps), and nearly every stage-2 list agrees.  A
    scale applied in the wrong place or not at all breaks this by orders of
    magnitude."""
    vol = _image(seed=13) if kind == "image" else _volume(seed=13)
    p = oracle.default_params(vol.ndim, vol.shape, nthreads=2)
    out1, pilot1, (s1, c1), (w1, wc1) = oracle.bmnd_poisson(vol, p, capture_matches=True)
    p.count_scale = 2.0
    out2, pilot2, (s2, c2), (w2, wc2) = oracle.bmnd_poisson(2.0 * vol, p, capture_matches=True)
    assert np.array_equal(c1, c2) and np.array_equal(s1, s2)
    np.testing.assert_allclose(pilot2, 2.0 * pilot1, rtol=2e-5, atol=1e-6)
    same = np.mean([np.array_equal(w1[g, :wc1[g]], w2[g, :wc2[g]]) for g in range(len(wc1))])
    assert same > 0.99
    np.testing.assert_allclose(out2, 2.0 * out1, rtol=1e-3, atol=1e-4)
    rel = float(np.max(np.abs(out2 - 2.0 * out1)) / np.max(np.abs(2.0 * out1)))
    assert rel < 1e-4


# ---- union-voxel entry order: numba's quicksort vs the stable order -------

def _numba_quicksort_argsort(a) -> np.ndarray:
    """numba's np.argsort (default kind 'quicksort', numba/misc/quicksort.py
    make_quicksort_impl with is_argsort=True): median-of-three partition with
    the pivot stashed at the end, insertion sort below 15 elements, the
    larger partition pushed on the stack.  Restated in Python."""
    A = list(a)
    R = list(range(len(A)))
    if len(A) < 2:
        return np.array(R)

    def insertion(low, high):
        for i in range(low + 1, high + 1):
            k, v, j = R[i], A[R[i]], i
            while j > low and v < A[R[j - 1]]:
                R[j] = R[j - 1]
                j -= 1
            R[j] = k

    def partition(low, high):
        mid = (low + high) >> 1
        if A[R[mid]] < A[R[low]]:
            R[low], R[mid] = R[mid], R[low]
        if A[R[high]] < A[R[mid]]:
            R[high], R[mid] = R[mid], R[high]
        if A[R[mid]] < A[R[low]]:
            R[low], R[mid] = R[mid], R[low]
        pivot = A[R[mid]]
        R[high], R[mid] = R[mid], R[high]
        i, j = low, high - 1
        while True:
            while i < high and A[R[i]] < pivot:
                i += 1
            while j >= low and pivot < A[R[j]]:
                j -= 1
            if i >= j:
                break
            R[i], R[j] = R[j], R[i]
            i += 1
            j -= 1
        R[i], R[high] = R[high], R[i]
        return i

    stack = [(0, len(A) - 1)]
    while stack:
        low, high = stack.pop()
        while high - low >= 15:
            i = partition(low, high)
            if high - i > i - low:
                if high > i:
                    stack.append((i + 1, high))
                high = i - 1
            else:
                if i > low:
                    stack.append((low, i - 1))
                low = i + 1
        insertion(low, high)
    return np.array(R)


def test_numba_argsort_restatements_agree():
    rng = np.random.default_rng(0)
    for n in (1, 2, 14, 15, 16, 64, 500, 4000):
        for hi in (3, 50, 10 ** 6):
            a = rng.integers(0, hi, n)
            py = _numba_quicksort_argsort(a)
            c = oracle.numba_argsort(a)
            assert np.array_equal(py, c)
            assert np.all(np.diff(a[py]) >= 0)
    # on few distinct labels the order inside equal labels is not the stable one
    a = rng.integers(0, 40, 4000)
    assert not np.array_equal(oracle.numba_argsort(a), np.argsort(a, kind="stable"))


def _ulps(a, b):
    ia = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    ib = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(ia - ib)


def test_quicksort_entry_order_deviation_is_a_few_ulps():
    """The reference sums a union voxel's entries in numba's argsort order;
    the oracle and the device in increasing member order.  Measured here:
    per-coefficient exact-plane variances differ by at most a few float32
    ulps and the pipeline outputs by well under the 1e-4 relative bar."""
    vol = _volume(seed=19, side=20)
    p = oracle.default_params(3, vol.shape, nthreads=2)
    block = tuple(p.wi_block[:3])
    counts = tuple(vol.shape[d] - block[d] + 1 for d in range(3))
    refs = oracle.schedule(counts, block, tuple(p.wi_step[:3]))
    var = _variance_model(vol, 1.0, 1e-10)
    m = oracle.blockmatch(refs, counts, tuple(p.wi_radius[:3]), p.wi_max, p.wi_min,
                          p.wi_threshold, volume=vol, block=block)
    fwd_s, _ = _spatial("dct", block)
    worst, changed = 0, 0
    try:
        for g in range(0, len(refs), 7):
            n = int(m.counts[g])
            apos = np.stack(np.unravel_index(m.selected[g, :n], counts), 1).astype(np.int64)
            G = oracle.haar(n)
            oracle.set_exact_order(False)
            stable = oracle.exact_variance(var, block, apos, G, fwd_s, min(4, n))
            oracle.set_exact_order(True)
            quick = oracle.exact_variance(var, block, apos, G, fwd_s, min(4, n))
            u = _ulps(stable, quick)
            worst = max(worst, int(u.max()))
            changed += int(np.count_nonzero(u))
        oracle.set_exact_order(True)
        out_q, pilot_q = oracle.bmnd_poisson(vol, p)
    finally:
        oracle.set_exact_order(False)
    out_s, pilot_s = oracle.bmnd_poisson(vol, p)
    rel = float(np.max(np.abs(out_q - out_s)) / np.max(np.abs(out_s)))
    print(f"entry order: {changed} coefficient variances moved, at most {worst} ulp; "
          f"pipeline output max rel {rel:.3g}")
    assert changed > 0  # the two orders do differ on real groups
    assert worst <= 8
    assert rel < 1e-5
