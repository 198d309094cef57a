"""Pins of the non-default aggregation weights and of soft thresholding in
the oracle, on the CPU:
- numpy's float64 sum (the group-scope risk) against np.sum;
- the diagonal-filter risks against the reference's dense-metric
  constructions (tests/test_aggregationweights_unit.py:20-225: the windowed
  group risk of every model against the diagonal of the dense Kronecker
  metric, the patch risks against dense patch-selection metrics, the
  unitary identities, classic proportional to variance, the extreme masks);
- soft thresholding against its closed form, and end to end against the
  reference's constant-patch test (tests/test_bmnd_smoke.py:49-70) restated
  for the Poisson model;
- the product's metric tables bit for bit equal to the oracle's, and its
  filter codes and option checks."""

import numpy as np
import pytest

import oracle
from bmnd_amd import BM3DProfile, BM4DProfile, NoiseModel
from bmnd_amd.engine import check_supported, filter_code, make_plan
from bmnd_amd.enums import AggregationWeightDomain, AggregationWeightModel, AggregationWeightScope


def _orthonormal(size, seed):
    q, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=(size, size)))
    return q.astype(np.float32)


def _metric_diagonals(inv_group, inv_spatial, window):
    s = inv_spatial.astype(np.float64)
    w = window.astype(np.float64)
    smd = np.diag(s.T @ ((w ** 2)[:, None] * s)).copy()
    g = inv_group.astype(np.float64)
    return smd, np.diag(g.T @ g).copy()


@pytest.mark.parametrize("n", [1, 2, 7, 8, 9, 15, 16, 17, 127, 128, 129, 255, 256, 257, 1000,
                               4000, 8191, 8192, 8193, 20000])
def test_float64_sum_follows_numpy(n):
    x = np.random.default_rng(n).normal(size=n) * 10.0 ** np.random.default_rng(n + 1).uniform(
        -8, 8, size=n)
    assert oracle.np_sum_f64(x) == float(np.sum(x, dtype=np.float64))


@pytest.mark.parametrize("model", ["classic", "variance", "risk"])
def test_group_risk_matches_the_dense_windowed_metric(model):
    n = 3
    inv_group = _orthonormal(n, 1)
    inv_spatial = np.array([[1.0, 0.3], [0.2, 0.8]], np.float32)
    window = np.array([0.5, 1.25], np.float32)
    att = np.array([[0.0, 0.2], [0.7, 1.0], [0.4, 0.9]], np.float32)
    var = np.array([[0.5, 0.7], [1.1, 1.3], [1.7, 1.9]], np.float32)
    sig = np.sqrt(var).astype(np.float32)
    signal = np.array([[2.0, 2.2], [2.4, 2.6], [2.8, 3.0]], np.float32)
    a = att.astype(np.float64)
    nv = (sig * sig).astype(np.float64)
    if model == "classic":
        err = a ** 2
    elif model == "variance":
        err = a ** 2 * nv
    else:
        err = a ** 2 * nv + (1.0 - a) ** 2 * signal
    synth = np.kron(inv_group.astype(np.float64), inv_spatial.astype(np.float64))
    rw = np.tile(window, n).astype(np.float64)
    metric = synth.T @ ((rw ** 2)[:, None] * synth)
    want = float(np.sum(np.diag(metric) * err.reshape(-1)))
    smd, gmd = _metric_diagonals(inv_group, inv_spatial, window)
    got = oracle.weight_risks(att, sig, inv_group, model=model, signal=signal, smd=smd, gmd=gmd,
                              windowed=True, group_scope=True)
    assert got.shape == (1,)
    assert got[0] == pytest.approx(want, rel=1e-12, abs=1e-12)


@pytest.mark.parametrize("windowed", [False, True])
def test_patch_risks_match_dense_patch_selection_metrics(windowed):
    n, p = 2, 3
    inv_group = _orthonormal(n, 2)
    inv_spatial = np.array([[1.0, 0.1, 0.0], [0.2, 0.8, 0.2], [0.0, 0.3, 0.7]], np.float32)
    window = np.array([0.4, 1.0, 0.7], np.float32)
    att = np.array([[0.2, 0.5, 0.8], [0.1, 0.6, 0.9]], np.float32)
    sig = np.sqrt(np.array([[0.7, 0.9, 1.1], [1.3, 1.5, 1.7]], np.float32)).astype(np.float32)
    err = att.astype(np.float64) ** 2 * (sig * sig).astype(np.float64)
    spatial = (inv_spatial.astype(np.float64).T
               @ ((window.astype(np.float64) ** 2)[:, None] * inv_spatial.astype(np.float64))
               if windowed else np.eye(p))
    want = np.empty(n)
    for i in range(n):
        row = inv_group[i].astype(np.float64)
        want[i] = np.sum(np.diag(np.kron(np.outer(row, row), spatial)) * err.reshape(-1))
    smd, _ = _metric_diagonals(inv_group, inv_spatial, window)
    got = oracle.weight_risks(att, sig, inv_group, smd=smd, windowed=windowed)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)


def test_unitary_identities():
    att = np.array([[0.2, 0.4], [0.6, 0.8]], np.float32)
    sig = np.sqrt(np.array([[0.5, 0.7], [1.1, 1.3]], np.float32)).astype(np.float32)
    ig = _orthonormal(2, 3)
    group = oracle.weight_risks(att, sig, ig, group_scope=True)[0]
    patch = oracle.weight_risks(att, sig, ig)
    assert np.sum(patch) == pytest.approx(group, rel=1e-7)
    # unitary synthesis and a constant window: windowed equals coefficient domain
    ig, isp = _orthonormal(2, 5), _orthonormal(2, 6)
    smd, gmd = _metric_diagonals(ig, isp, np.ones(2, np.float32))
    for scope in (False, True):
        coef = oracle.weight_risks(att, sig, ig, group_scope=scope)
        win = oracle.weight_risks(att, sig, ig, smd=smd, gmd=gmd, windowed=True, group_scope=scope)
        assert np.allclose(win, coef, rtol=1e-7, atol=1e-9)


def test_classic_is_variance_over_a_constant_variance():
    att = (np.arange(1, 7, dtype=np.float32).reshape(2, 3) / 7.0).astype(np.float32)
    sig = np.full_like(att, np.float32(np.sqrt(2.5)))
    ig = np.eye(2, dtype=np.float32)
    classic = oracle.weight_risks(att, sig, ig, model="classic", group_scope=True)[0]
    variance = oracle.weight_risks(att, sig, ig, group_scope=True)[0]
    assert variance == pytest.approx(float(sig[0, 0]) ** 2 * classic)


@pytest.mark.parametrize("retained", [False, True])
def test_extreme_masks(retained):
    att = np.full((2, 3), float(retained), np.float32)
    var = np.arange(1, 7, dtype=np.float32).reshape(2, 3)
    sig = np.sqrt(var).astype(np.float32)
    risk = oracle.weight_risks(att, sig, np.eye(2, dtype=np.float32), group_scope=True)[0]
    nv = (sig * sig).astype(np.float64)
    assert risk == (float(np.sum(nv, dtype=np.float64)) if retained else 0.0)


def test_risk_adds_the_rejected_signal():
    rng = np.random.default_rng(4)
    att = rng.uniform(0, 1, (4, 6)).astype(np.float32)
    sig = rng.uniform(0.5, 2, (4, 6)).astype(np.float32)
    signal = rng.uniform(0, 5, (4, 6)).astype(np.float32)
    ig = oracle.group_inverse(4)
    v = oracle.weight_risks(att, sig, ig)
    r = oracle.weight_risks(att, sig, ig, model="risk", signal=signal)
    extra = ((1.0 - att.astype(np.float64)) ** 2 * signal).sum(axis=1)
    assert np.allclose(r - v, (ig.astype(np.float64) ** 2) @ extra, rtol=1e-12)


def test_soft_threshold_closed_form():
    c = np.array([3.0, -3.0, 0.5, -0.5, 0.0, -0.0, 2.0, -2.0], np.float32)
    s = np.array([1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 0.5, 0.5], np.float32)
    out, att = oracle.ht_soft(c, s, lam=1.0)
    shr = np.maximum(np.abs(c) - np.float32(1.0) * s, np.float32(0.0))
    want = (np.sign(c) * shr).astype(np.float32)
    want_att = np.divide(shr, np.abs(c), out=np.zeros_like(c), where=np.abs(c) > 0)
    assert np.array_equal(out.view(np.int32), want.view(np.int32))
    assert np.array_equal(att, want_att)
    assert list(out[:4]) == [2.0, -2.0, 0.0, 0.0] and out[3] == 0.0 and np.signbit(out[3])


def test_soft_threshold_shrinks_the_dc_of_a_constant_image():
    """tests/test_bmnd_smoke.py:49-70 restated for the Poisson model: an 8x8
    constant image of 2 counts, single-patch groups (radius 0, stack 1), k =
    0 and the orthonormal DCT in stage 1; every coefficient's std is sqrt(2)
    (orthonormal rows, the variance model of constant counts), so the DC
    8 * 2 shrinks by lambda * sqrt(2) and every other coefficient is 0: the
    pilot is 2 - lambda sqrt(2) / 8 everywhere."""
    img = np.full((8, 8), 2.0, np.float32)
    p = oracle.default_params(2, img.shape)
    p.ht_radius[0] = p.ht_radius[1] = 0
    p.ht_max = p.ht_min = 1
    p.k = 0
    p.ht_soft = 1
    p._keep = (oracle.reference_spatial_matrices("dct", (8, 8)), p._keep[1])
    p.ht_spatial = p._keep[0].ctypes.data
    _, pilot = oracle.bmnd_poisson(img, p)
    assert np.allclose(pilot, 2.0 - 3.0 * np.sqrt(2.0) / 8.0, atol=1e-6)
    p.ht_soft = 0
    _, hard = oracle.bmnd_poisson(img, p)
    assert np.allclose(hard, 2.0, atol=1e-6)


@pytest.mark.parametrize("kind", ["image", "volume"])
def test_product_metric_tables_equal_the_oracle(kind):
    prof = (BM3DProfile if kind == "image" else BM4DProfile)(noise_model=NoiseModel.POISSON)
    ndim = 2 if kind == "image" else 3
    p = oracle.default_params(ndim, (40,) * ndim)
    for stage in ("ht", "wiener"):
        plan = make_plan(prof, stage, ndim)
        table = plan.matrix_table()
        P = plan.P
        m = plan.max_m
        gtot = m * (m + 1) * (2 * m + 1) // 6
        sf = sum(b * b for b in plan.block3)
        off = 2 * gtot + 2 * sf
        off += off % 2
        n64 = P + m * (m + 1) + sum(plan.block3)
        f64 = table[off:off + 2 * n64].view(np.float64)
        smd, gmd = oracle.synthesis_metric_tables(p, stage)
        assert np.array_equal(f64[:P], smd)
        assert np.array_equal(f64[P:P + gmd.size], gmd)
        # then the pair tables of the pair-order exact planes, per axis and
        # offset e: H[t][p] = A[t][p] A[t][p - e + b - 1], 0 off the block
        o = off + 2 * n64
        fwd = table[2 * gtot:2 * gtot + sf]
        fo = 0
        for b in plan.block3:
            A = fwd[fo:fo + b * b].reshape(b, b)
            fo += b * b
            h = table[o:o + (2 * b - 1) * b * b].reshape(2 * b - 1, b, b)
            o += (2 * b - 1) * b * b
            for e in range(2 * b - 1):
                d = e - (b - 1)
                for t in range(b):
                    for q in range(b):
                        want = A[t, q] * A[t, q - d] if 0 <= q - d < b else np.float32(0)
                        assert h[e, t, q] == want
        # then the spatial metric SM = Sinv^T diag(w^2) Sinv of the exact group
        # risk (variance.py:1543-1556), from the synthesis matrices and window
        inv = table[2 * gtot + sf:2 * gtot + 2 * sf]
        Si, io = np.ones((1, 1)), 0
        for b in plan.block3:
            Si = np.kron(Si, inv[io:io + b * b].reshape(b, b).astype(np.float64))
            io += b * b
        w2 = (plan.kaiser.reshape(-1).astype(np.float32) ** 2).astype(np.float64)
        sm = table[o:o + P * P].reshape(P, P)
        assert np.allclose(sm, Si.T @ (w2[:, None] * Si), rtol=1e-6, atol=1e-7)
        assert np.array_equal(np.diag(sm), np.diag(sm.astype(np.float64)).astype(np.float32))
        assert o + P * P == table.size


def test_filter_codes_and_option_checks():
    prof = BM4DProfile(noise_model=NoiseModel.POISSON)
    assert filter_code(prof, "ht", 1) == 0 and filter_code(prof, "wiener", 1) == 1
    prof.ht_use_soft_thresholding = True
    prof.ht_weight_model = AggregationWeightModel.CLASSIC
    prof.wiener_weight_model = AggregationWeightModel.RISK
    prof.wiener_weight_domain = AggregationWeightDomain.WINDOWED_SYNTHESIS
    prof.wiener_weight_scope = AggregationWeightScope.PATCH
    prof.ht_weight_scope = AggregationWeightScope.GROUP
    assert filter_code(prof, "ht", 1) == 4 | (1 << 3) | 64
    assert filter_code(prof, "wiener", 3) == 3 | (2 << 3) | 32
    assert filter_code(prof, "wiener", 3, coefficients=True) == 3 | (2 << 3) | 32 | 128

    class R:
        is_poisson = True
        from bmnd_amd.enums import BlockMatchDistance
        blockmatch_distance = BlockMatchDistance.POISSON_DEVIANCE
        variance_k = 4
    check_supported(prof, R())
    # the exact covariance-aware group risk runs under the Poisson model and
    # is refused under the Gaussian one
    prof.wiener_weight_scope = AggregationWeightScope.GROUP
    check_supported(prof, R())
    R.is_poisson = False
    R.blockmatch_distance = R.BlockMatchDistance.SSD
    with pytest.raises(NotImplementedError, match="exact covariance-aware"):
        check_supported(prof, R())
    R.variance_k = 0
    check_supported(prof, R())
