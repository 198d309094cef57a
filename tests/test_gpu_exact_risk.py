"""The exact covariance-aware group risk (wiener.py:638-844) on the GPU
against the oracle with the same options.  The device regroups the union-
voxel covariance by member pairs and transforms it separably, so the weights
agree with the oracle's to float32 rounding, not bit for bit: a stage whose
weights take the exact risk is held to 1e-4 max relative error (the bar of
the float32 path); everything before it stays bit-identical (the HT lists,
and the pilot when only the Wiener stage uses the exact risk)."""

import numpy as np
import pytest

import oracle
from bmnd_amd import bmnd, bmnd_batch

from test_gpu_parity import _compare
from test_gpu_weights import _configure, _input

pytestmark = pytest.mark.gpu

# (HT model, windowed, group), (Wiener model, windowed, group), soft, k
CASES = [
    (("variance", False, False), ("variance", True, True), False, 4),
    (("variance", False, False), ("risk", True, True), False, 4),
    (("variance", True, True), ("variance", False, False), False, 4),
    (("variance", True, True), ("risk", True, True), True, 2),
]


@pytest.mark.parametrize("kind", ["volume", "image"])
@pytest.mark.parametrize("case", CASES, ids=[f"case{i}" for i in range(len(CASES))])
def test_exact_group_risk_matches_oracle(kind, case):
    (hm, hw, hg), (wm, ww, wg), soft, k = case
    print(f"exact risk: {kind} ht {hm}/{hw}/{hg} wiener {wm}/{ww}/{wg} soft {soft} k {k}",
          flush=True)
    noisy, prof = _input(kind)
    params = oracle.default_params(noisy.ndim, noisy.shape)
    prof.k = params.k = k
    prof.ht_use_soft_thresholding = soft
    params.ht_soft = int(soft)
    _configure(prof, params, "ht", hm, hw, hg)
    _configure(prof, params, "wiener", wm, ww, wg)
    stats, out, o_out = _compare(noisy, prof, params, "auto-fast")
    assert stats["ht_groups_differing"] == 0
    ht_exact = hw and hg
    if not ht_exact:
        assert stats["pilot_bits_differing"] == 0
        assert stats["wiener_groups_differing"] == 0
    rel = float(np.max(np.abs(out - o_out)) / np.max(np.abs(o_out)))
    assert rel <= 1e-4, rel


def test_batch_runs_exact_risk_input_by_input():
    noisy, prof = _input("volume")
    params = oracle.default_params(noisy.ndim, noisy.shape)
    _configure(prof, params, "wiener", "variance", True, True)
    vols = [noisy, np.ascontiguousarray(noisy[:20, 3:25, :18])]
    print("exact risk: batch of two volumes", flush=True)
    outs = bmnd_batch(vols, prof)
    for v, o in zip(vols, outs):
        assert np.array_equal(o.view(np.int32), bmnd(v, prof).view(np.int32))
