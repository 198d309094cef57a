"""HT soft thresholding (bmndalgo.py:475-483) and the non-default
aggregation weights (aggregationweights.py:123-259: the classic and risk
models, the windowed-synthesis domain, the group scope) on the GPU against
the CPU oracle with the same options, bit for bit in both stages' lists, the
pilot and the output; bmnd_batch under them equal to bmnd() per input."""

import numpy as np
import pytest
import torch

import oracle
from bmnd_amd import BM3DProfile, BM4DProfile, NoiseModel, bmnd, bmnd_batch, phantoms
from bmnd_amd.enums import AggregationWeightDomain, AggregationWeightModel, AggregationWeightScope

from test_gpu_parity import _compare

pytestmark = pytest.mark.gpu

M = {"variance": AggregationWeightModel.VARIANCE, "classic": AggregationWeightModel.CLASSIC,
     "risk": AggregationWeightModel.RISK}


def _input(kind):
    if kind == "volume":
        _, noisy = phantoms.ellipsoid_volume((24, 26, 22), seed=81, peak=10.0, n_ellipsoids=6,
                                             axes_range=(3.0, 10.0))
        return noisy, BM4DProfile(noise_model=NoiseModel.POISSON)
    _, noisy = phantoms.blob_image(64, seed=82, peak=20.0)
    return noisy, BM3DProfile(noise_model=NoiseModel.POISSON)


def _configure(prof, params, stage, model, windowed, group):
    setattr(prof, f"{stage}_weight_model", M[model])
    setattr(prof, f"{stage}_weight_domain", AggregationWeightDomain.WINDOWED_SYNTHESIS if windowed
            else AggregationWeightDomain.COEFFICIENT)
    setattr(prof, f"{stage}_weight_scope", AggregationWeightScope.GROUP if group
            else AggregationWeightScope.PATCH)
    oracle.set_weight_options(params, stage, model, windowed, group)


# (HT model, windowed, group), (Wiener model, windowed, group), soft, k
CASES = [
    (("variance", False, False), ("variance", False, False), True, 4),
    (("classic", False, False), ("classic", False, False), False, 4),
    (("variance", True, False), ("risk", False, False), False, 4),
    (("variance", False, True), ("risk", True, False), True, 4),
    (("classic", True, True), ("classic", True, True), False, 4),
    (("variance", True, True), ("risk", True, True), False, 0),
    (("classic", False, True), ("variance", False, True), True, 2),
]


@pytest.mark.parametrize("kind", ["volume", "image"])
@pytest.mark.parametrize("case", CASES, ids=[f"case{i}" for i in range(len(CASES))])
def test_weights_and_soft_threshold_match_oracle(kind, case):
    (hm, hw, hg), (wm, ww, wg), soft, k = case
    print(f"weights: {kind} ht {hm}/{hw}/{hg} wiener {wm}/{ww}/{wg} soft {soft} k {k}", flush=True)
    noisy, prof = _input(kind)
    params = oracle.default_params(noisy.ndim, noisy.shape)
    prof.k = params.k = k
    prof.ht_use_soft_thresholding = soft
    params.ht_soft = int(soft)
    _configure(prof, params, "ht", hm, hw, hg)
    _configure(prof, params, "wiener", wm, ww, wg)
    stats, _, _ = _compare(noisy, prof, params, "auto-fast")
    assert stats["ht_groups_differing"] == 0
    assert stats["pilot_bits_differing"] == 0
    assert stats["wiener_groups_differing"] == 0
    assert stats["out_bits_differing"] == 0


def test_options_change_the_estimate_and_batch_matches_bmnd():
    noisy, prof = _input("volume")
    base = bmnd(noisy, prof)
    prof.ht_use_soft_thresholding = True
    prof.wiener_weight_model = AggregationWeightModel.RISK
    prof.wiener_weight_domain = AggregationWeightDomain.WINDOWED_SYNTHESIS
    out = bmnd(noisy, prof)
    assert not np.array_equal(out, base)
    vols = [noisy, np.ascontiguousarray(noisy[:20, 3:25, :18]), np.ascontiguousarray(noisy[2:, :, 4:])]
    print("weights: batch of three volumes", flush=True)
    outs = bmnd_batch(vols, prof)
    for v, o in zip(vols, outs):
        assert np.array_equal(o.view(np.int32), bmnd(v, prof).view(np.int32))
