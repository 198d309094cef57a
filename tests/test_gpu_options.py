"""Non-default Poisson options on the GPU against the CPU oracle, bit for
bit: the Wiener stage matching on the observation instead of the pilot
(poisson_wiener_match_source, bmndalgo.py:714-720; its counts then use the
count-table matcher) and taking the observation's variance model
(poisson_variance_source_wiener, bmndalgo.py:635-646), alone and together;
and the per-reference stage-1 thresholds of poisson_ht_match_policy=
"reference_finite_count" (bmndalgo.py:269-284) in every matching mode and
through bmnd_batch."""

import numpy as np
import pytest
import torch

import oracle
from bmnd_amd import BM3DProfile, BM4DProfile, NoiseModel, bmnd, bmnd_batch, phantoms
from bmnd_amd.engine import PoissonPipeline
from bmnd_amd.enums import PoissonHTMatchPolicy, PoissonVarianceSource

from test_gpu_parity import _compare

pytestmark = pytest.mark.gpu

OBS, PILOT = PoissonVarianceSource.OBSERVATION, PoissonVarianceSource.PILOT


@pytest.mark.parametrize("match_src,var_src", [(OBS, PILOT), (PILOT, OBS), (OBS, OBS)])
@pytest.mark.parametrize("kind", ["volume", "image"])
def test_wiener_sources_match_oracle(kind, match_src, var_src):
    print(f"options: {kind}, match {match_src.value}, variance {var_src.value}", flush=True)
    if kind == "volume":
        _, noisy = phantoms.ellipsoid_volume((26, 28, 24), seed=71, peak=10.0, n_ellipsoids=6,
                                             axes_range=(3.0, 10.0))
        prof = BM4DProfile(noise_model=NoiseModel.POISSON)
    else:
        _, noisy = phantoms.blob_image(72, seed=72, peak=20.0)
        prof = BM3DProfile(noise_model=NoiseModel.POISSON)
    prof.poisson_wiener_match_source = match_src
    prof.poisson_variance_source_wiener = var_src
    params = oracle.default_params(noisy.ndim, noisy.shape)
    params.wi_match_observation = int(match_src is OBS)
    params.wi_var_observation = int(var_src is OBS)
    stats, _, _ = _compare(noisy, prof, params, "auto-fast")
    assert stats["ht_groups_differing"] == 0
    assert stats["wiener_groups_differing"] == 0
    assert stats["pilot_bits_differing"] == 0
    assert stats["out_bits_differing"] == 0
    if match_src is OBS:
        assert stats["modes"]["wiener"] == 2  # integer counts: the count table


def _policy_input(kind):
    if kind.endswith("volume"):
        _, noisy = phantoms.ellipsoid_volume((24, 26, 22), seed=73, peak=10.0, n_ellipsoids=6,
                                             axes_range=(3.0, 10.0))
        prof = BM4DProfile(noise_model=NoiseModel.POISSON)
    else:
        # piecewise-constant rates 10..50: thresholds bind in most windows
        rng = np.random.default_rng(74)
        rate = np.kron(rng.integers(1, 6, (8, 8)), np.ones((8, 8))) * 10.0
        noisy = np.ascontiguousarray(rng.poisson(rate).astype(np.float32)[:57, :64])
        prof = BM3DProfile(noise_model=NoiseModel.POISSON)
    if kind.startswith("scaled"):
        # non-integer counts: the float-pass / certified / exact matchers
        noisy = (noisy * np.float32(0.37)).astype(np.float32)
    return noisy, prof


def _with_policy(prof, beta, power):
    prof.poisson_ht_match_policy = PoissonHTMatchPolicy.REFERENCE_FINITE_COUNT
    prof.poisson_ht_match_structure_beta = beta
    prof.poisson_ht_match_intensity_power = power
    return prof


@pytest.mark.parametrize("kind,deviance,lam,beta,power", [
    ("volume", "auto-fast", 2.9527, 0.1, 1.0), ("volume", "auto-fast", 1.0, 0.4, 0.5),
    ("image", "auto-fast", 2.9527, 0.4, 0.5), ("scaled-image", "auto", 2.9527, 0.1, 1.0),
    ("scaled-image", "auto-certified", 1.0, 0.1, 1.0),
    ("scaled-image", "auto-fast", 1.0, 0.4, 0.5), ("scaled-volume", "auto-fast", 0.5, 0.1, 1.0)])
def test_reference_finite_count_matches_oracle(kind, deviance, lam, beta, power):
    """Pilot, output and match lists against the oracle pipeline fed the
    oracle's thresholds; stage-1 accepted counts (every reference's
    threshold decision over its whole window) against the oracle matcher."""
    print(f"reference_finite_count: {kind}, {deviance}, lambda {lam}, beta {beta}, "
          f"power {power}", flush=True)
    noisy, prof = _policy_input(kind)
    prof = _with_policy(prof, beta, power)
    prof.ht_match_threshold = lam
    params = oracle.default_params(noisy.ndim, noisy.shape)
    thr = oracle.poisson_reference_thresholds(noisy, params, lam, beta, power)
    oracle.set_ht_thresholds(params, thr)
    stats, _, _ = _compare(noisy, prof, params, deviance)
    assert stats["ht_groups_differing"] == 0
    assert stats["wiener_groups_differing"] == 0
    assert stats["pilot_bits_differing"] == 0
    assert stats["out_bits_differing"] == 0
    dev = torch.device("cuda", 0)
    pipe = PoissonPipeline(prof, noisy.shape, dev, deviance=deviance)
    ht, _ = pipe.run(torch.from_numpy(noisy).to(dev))
    plan = pipe.plans["ht"]
    counts = [noisy.shape[d] - plan.block[d] + 1 for d in range(noisy.ndim)]
    refs = plan.schedule.reference_abs
    radius = tuple(int(r) for r in prof.ht_search_window)
    o = oracle.blockmatch(refs, counts, radius, plan.max_m, plan.min_m, thr, volume=noisy,
                          block=plan.block)
    acc = ht.accepted.cpu().numpy()
    assert np.array_equal(acc, o.accepted)
    assert np.array_equal(ht.counts.cpu().numpy(), o.counts)
    print(f"accepted counts: min {acc.min()} median {np.median(acc)} max {acc.max()}", flush=True)


def test_reference_finite_count_batch_matches_bmnd():
    """bmnd_batch under the policy: integer and non-integer images of
    uneven shapes, each equal to bmnd() alone (and one to the oracle)."""
    prof = _with_policy(BM3DProfile(noise_model=NoiseModel.POISSON), 0.1, 1.0)
    imgs = []
    for side, seed, peak in (((40, 45), 75, 5.0), ((64, 33), 76, 30.0), ((19, 50), 77, 10.0)):
        _, img = phantoms.blob_image(max(max(side), 32), seed=seed, peak=peak)
        imgs.append(np.ascontiguousarray(img[:side[0], :side[1]]))
    imgs[2] = (imgs[2] * np.float32(0.37)).astype(np.float32)
    print("reference_finite_count: bmnd_batch of 3 images", flush=True)
    outs = bmnd_batch(imgs, prof)
    for i, x in enumerate(imgs):
        assert np.array_equal(outs[i].view(np.int32), bmnd(x, prof).view(np.int32)), f"input {i}"
    params = oracle.default_params(2, imgs[0].shape)
    oracle.set_ht_thresholds(params, oracle.poisson_reference_thresholds(
        imgs[0], params, prof.ht_match_threshold, 0.1, 1.0))
    ref, _ = oracle.bmnd_poisson(imgs[0], params)
    assert np.array_equal(outs[0].view(np.int32), ref.view(np.int32))


_CODES = {"poisson_deviance": oracle.DIST_DEVIANCE, "pearson": oracle.DIST_PEARSON,
          "anscombe_ssd": oracle.DIST_ANSCOMBE}


@pytest.mark.parametrize("kind,dist,policy", [
    ("volume", "pearson", "fixed"), ("image", "anscombe_ssd", "fixed"),
    ("scaled-volume", "anscombe_ssd", "fixed"), ("scaled-image", "pearson", "fixed"),
    ("volume", "poisson_deviance", "candidate_standardized"),
    ("image", "pearson", "candidate_standardized"),
    ("scaled-image", "anscombe_ssd", "candidate_standardized"),
    ("volume", "anscombe_ssd", "reference_finite_count")])
def test_distances_and_standardized_policy_match_oracle(kind, dist, policy):
    """The Pearson and Anscombe SSD distances in both stages and the
    candidate-standardized stage-1 distance (blockmatching.py:322-426):
    lists, pilot and output bit-identical to the oracle; stage-1 accepted
    counts equal to the oracle matcher's."""
    print(f"distance: {kind}, {dist}, policy {policy}", flush=True)
    noisy, prof = _policy_input(kind)
    prof.blockmatch_distance = dist
    prof.poisson_ht_match_policy = policy
    params = oracle.default_params(noisy.ndim, noisy.shape)
    params.distance = _CODES[dist]
    thr = np.float32(prof.ht_match_threshold)
    null = None
    if policy == "candidate_standardized":
        oracle.set_ht_standardized(params, 512)
        null = params._null_keep
    elif policy == "reference_finite_count":
        thr = oracle.poisson_reference_thresholds(noisy, params, prof.ht_match_threshold,
                                                  prof.poisson_ht_match_structure_beta,
                                                  prof.poisson_ht_match_intensity_power)
        oracle.set_ht_thresholds(params, thr)
    stats, _, _ = _compare(noisy, prof, params, "auto-fast")
    assert stats["ht_groups_differing"] == 0
    assert stats["wiener_groups_differing"] == 0
    assert stats["pilot_bits_differing"] == 0
    assert stats["out_bits_differing"] == 0
    dev = torch.device("cuda", 0)
    pipe = PoissonPipeline(prof, noisy.shape, dev)
    ht, _ = pipe.run(torch.from_numpy(noisy).to(dev))
    plan = pipe.plans["ht"]
    counts = [noisy.shape[d] - plan.block[d] + 1 for d in range(noisy.ndim)]
    o = oracle.blockmatch(plan.schedule.reference_abs, counts,
                          tuple(int(r) for r in prof.ht_search_window), plan.max_m, plan.min_m,
                          thr, volume=noisy, block=plan.block, code=_CODES[dist],
                          null_moments=null)
    acc = ht.accepted.cpu().numpy()
    assert np.array_equal(acc, o.accepted)
    assert np.array_equal(ht.counts.cpu().numpy(), o.counts)
    print(f"accepted counts: min {acc.min()} median {np.median(acc)} max {acc.max()}", flush=True)


def test_standardized_pearson_batch_matches_bmnd():
    prof = BM3DProfile(noise_model=NoiseModel.POISSON)
    prof.blockmatch_distance = "pearson"
    prof.poisson_ht_match_policy = "candidate_standardized"
    imgs = []
    for side, seed, peak in (((40, 45), 78, 5.0), ((64, 33), 79, 30.0), ((19, 50), 80, 10.0)):
        _, img = phantoms.blob_image(max(max(side), 32), seed=seed, peak=peak)
        imgs.append(np.ascontiguousarray(img[:side[0], :side[1]]))
    imgs[1] = (imgs[1] * np.float32(0.37)).astype(np.float32)
    print("candidate_standardized pearson: bmnd_batch of 3 images", flush=True)
    outs = bmnd_batch(imgs, prof)
    for i, x in enumerate(imgs):
        assert np.array_equal(outs[i].view(np.int32), bmnd(x, prof).view(np.int32)), f"input {i}"
    params = oracle.default_params(2, imgs[0].shape)
    params.distance = oracle.DIST_PEARSON
    oracle.set_ht_standardized(params, 512)
    ref, _ = oracle.bmnd_poisson(imgs[0], params)
    assert np.array_equal(outs[0].view(np.int32), ref.view(np.int32))


@pytest.mark.parametrize("kind,mode", [("volume", "classic"), ("image", "noise-floor"),
                                       ("scaled-volume", "noise-floor"),
                                       ("scaled-image", "classic")])
def test_wiener_gain_modes_match_oracle(kind, mode):
    """poisson_wiener_gain_mode 'classic' and 'noise-floor' (wiener.py:24-53):
    pilot, output and lists bit-identical to the oracle."""
    print(f"gain mode: {kind}, {mode}", flush=True)
    noisy, prof = _policy_input(kind)
    prof.poisson_wiener_gain_mode = mode
    params = oracle.default_params(noisy.ndim, noisy.shape)
    params.wi_gain_mode = oracle.GAIN_MODES[mode]
    stats, out, _ = _compare(noisy, prof, params, "auto-fast")
    assert stats["ht_groups_differing"] == 0
    assert stats["wiener_groups_differing"] == 0
    assert stats["pilot_bits_differing"] == 0
    assert stats["out_bits_differing"] == 0
    default, _ = oracle.bmnd_poisson(noisy, oracle.default_params(noisy.ndim, noisy.shape))
    assert not np.array_equal(out, default)  # the option changes the estimate


@pytest.mark.parametrize("kind,scale,dist,policy", [
    ("scaled-volume", 2.5, "pearson", "candidate_standardized"),
    ("image", 0.5, "poisson_deviance", "reference_finite_count"),
    ("volume", 1.7, "anscombe_ssd", "fixed")])
def test_options_with_count_scale_match_oracle(kind, scale, dist, policy):
    """The new distances and policies at a count scale other than 1 (counts
    divided by poisson_count_scale in every distance and table lookup)."""
    print(f"count scale {scale}: {kind}, {dist}, {policy}", flush=True)
    noisy, prof = _policy_input(kind)
    prof.blockmatch_distance = dist
    prof.poisson_ht_match_policy = policy
    prof.poisson_count_scale = scale
    params = oracle.default_params(noisy.ndim, noisy.shape)
    params.distance = _CODES[dist]
    params.count_scale = scale
    if policy == "candidate_standardized":
        oracle.set_ht_standardized(params, 512)
    elif policy == "reference_finite_count":
        oracle.set_ht_thresholds(params, oracle.poisson_reference_thresholds(
            noisy, params, prof.ht_match_threshold, prof.poisson_ht_match_structure_beta,
            prof.poisson_ht_match_intensity_power))
    stats, _, _ = _compare(noisy, prof, params, "auto-fast")
    assert stats["ht_groups_differing"] == 0
    assert stats["wiener_groups_differing"] == 0
    assert stats["pilot_bits_differing"] == 0
    assert stats["out_bits_differing"] == 0


def test_sigma_maps_under_options_match_oracle():
    """return_sigma_map with the standardized Anscombe policy and the
    noise-floor gain: estimate and both maps bit-identical to the oracle."""
    noisy, prof = _policy_input("volume")
    prof.blockmatch_distance = "anscombe_ssd"
    prof.poisson_ht_match_policy = "candidate_standardized"
    prof.poisson_wiener_gain_mode = "noise-floor"
    print("sigma maps under options", flush=True)
    out, maps = bmnd(noisy, prof, return_sigma_map=True)
    params = oracle.default_params(3, noisy.shape)
    params.distance = oracle.DIST_ANSCOMBE
    params.wi_gain_mode = oracle.GAIN_MODES["noise-floor"]
    oracle.set_ht_standardized(params, 512)
    o_out, o_maps = oracle.bmnd_poisson_sigma_maps(noisy, params)
    assert np.array_equal(out.view(np.int32), o_out.view(np.int32))
    assert maps["global"] == o_maps["global"]
    for stage in ("ht", "wiener"):
        assert np.array_equal(maps[stage].view(np.int32), o_maps[stage].view(np.int32)), stage


def test_finite_count_anscombe_volume_batch_matches_bmnd():
    """bmnd_batch over uneven volumes with the Anscombe SSD distance and
    per-reference thresholds: each input equal to bmnd() alone."""
    prof = BM4DProfile(noise_model=NoiseModel.POISSON)
    prof.blockmatch_distance = "anscombe_ssd"
    prof.poisson_ht_match_policy = "reference_finite_count"
    vols = []
    for shape, seed in (((20, 24, 18), 91), ((16, 16, 30), 92), ((9, 28, 22), 93)):
        _, v = phantoms.ellipsoid_volume(shape, seed=seed, peak=8.0, n_ellipsoids=4,
                                         axes_range=(2.0, 8.0))
        vols.append(v)
    vols[1] = (vols[1] * np.float32(0.37)).astype(np.float32)
    print("reference_finite_count anscombe: bmnd_batch of 3 volumes", flush=True)
    outs = bmnd_batch(vols, prof)
    for i, x in enumerate(vols):
        assert np.array_equal(outs[i].view(np.int32), bmnd(x, prof).view(np.int32)), f"input {i}"
