"""Host-side behaviour of the drop-in API that needs no GPU: profile
defaults, validation order and messages (tests/test_bmnd_pipeline_unit.py of
the reference), enum coercion and the all-zero fast path."""

import numpy as np
import pytest

from bmnd_amd import (BM3DProfile, BM4DProfile, BlockMatchDistance, NoiseModel,
                      PoissonWienerGainMode, ReferenceScheduleMode, bmnd)
from bmnd_amd.profilevalidation import resolve_profile


def _poisson(profile_cls=BM3DProfile, **kw):
    return profile_cls(noise_model=NoiseModel.POISSON, **kw)


def test_profile_defaults():
    p3, p4 = BM3DProfile(), BM4DProfile()
    assert (p3.ht_block_size, p3.wiener_block_size, p3.ht_search_window) == ((8, 8), (8, 8), (19, 19))
    assert (p4.ht_block_size, p4.wiener_block_size, p4.ht_search_window) == ((4, 4, 4), (5, 5, 5), (7, 7, 7))
    for p in (p3, p4):
        assert (p.ht_step, p.ht_max_stack_size, p.wiener_max_stack_size) == ((3,) * len(p.ht_step), 16, 32)
        assert p.ht_match_threshold == 2.9527 and p.ht_lambda_threshold == 3.0
        assert p.wiener_variance_scale == 0.4 and p.k == 4
        assert p.reference_schedule_mode is ReferenceScheduleMode.GENERATED
    assert p3.wiener_match_threshold == 0.3937 and p4.wiener_match_threshold == 0.7689


def test_distance_resolves_to_deviance_and_strings_are_coerced():
    p = BM3DProfile(noise_model="poisson", poisson_wiener_gain_mode="variance-scaled")
    r = resolve_profile(p, (64, 64))
    assert r.is_poisson and r.blockmatch_distance is BlockMatchDistance.POISSON_DEVIANCE
    assert p.noise_model is NoiseModel.POISSON
    assert p.poisson_wiener_gain_mode is PoissonWienerGainMode.VARIANCE_SCALED
    assert r.variance_k == 4 and p.k == 4 and p.gamma == 3.0


def test_sigma_and_sigma_psd_are_exclusive():
    with pytest.raises(ValueError, match="mutually exclusive"):
        bmnd(np.ones((16, 16)), _poisson(), sigma=1.0, sigma_psd=1.0)


def test_poisson_rejects_sigma_psd():
    with pytest.raises(ValueError, match="sigma_psd is only supported"):
        bmnd(np.ones((16, 16)), _poisson(), sigma_psd=1.0)


def test_rejects_negative_and_non_finite_input():
    with pytest.raises(ValueError, match="non-negative"):
        bmnd(-np.ones((16, 16)), _poisson())
    x = np.ones((16, 16))
    x[3, 3] = np.nan
    with pytest.raises(ValueError, match="finite values"):
        bmnd(x, _poisson())
    with pytest.raises(ValueError, match="representable as finite float32"):
        bmnd(np.full((16, 16), 1e300), _poisson())


@pytest.mark.parametrize("field,value,msg", [
    ("ht_block_size", (8,), "must contain 2 values"),
    ("ht_step", (0, 3), "must be positive"),
    ("ht_min_stack_size", 40, "cannot exceed"),
    ("ht_match_threshold", -1.0, "finite and non-negative"),
    ("k", -1, "non-negative integer"),
    ("poisson_count_scale", 0.0, "finite and positive"),
    ("ref_batch_size", 0, "positive integer"),
])
def test_profile_validation_messages(field, value, msg):
    p = _poisson()
    setattr(p, field, value)
    with pytest.raises(ValueError, match=msg):
        bmnd(np.ones((32, 32)), p)


def test_cache_argument_validation():
    with pytest.raises(ValueError, match="max_cache_bytes"):
        bmnd(np.zeros((16, 16)), _poisson(), max_cache_bytes=-3)


def test_all_zero_volume_fast_path_runs_no_kernel():
    calls = []
    z = np.zeros((20, 20), np.float32)
    out = bmnd(z, _poisson(), stage_callback=lambda name, est: calls.append((name, est.shape)))
    assert out is z
    assert calls == [("ht", (20, 20)), ("wiener", (20, 20))]
    out, maps = bmnd(z, _poisson(), return_sigma_map=True)
    assert set(maps) == {"ht", "wiener"}


@pytest.mark.parametrize("fields", [
    {"blockmatch_distance": "poisson_deviance"},
    {"ht_weight_domain": "windowed_synthesis", "ht_weight_scope": "group"},
])
def test_unsupported_options_raise_before_any_gpu_work(fields):
    # options the Gaussian model does not run on the device
    p = BM3DProfile()
    for field, value in fields.items():
        setattr(p, field, value)
    with pytest.raises(NotImplementedError):
        bmnd(np.ones((32, 32)), p, sigma=0.1)


def test_poisson_ssd_is_supported():
    # the SSD distance under the Poisson model runs on the device (stage 1
    # with the Poisson gamma correction, bmndalgo.py:292-294, 321-322)
    from bmnd_amd.engine import check_supported
    from bmnd_amd.profilevalidation import resolve_profile
    p = _poisson()
    p.blockmatch_distance = "ssd"
    check_supported(p, resolve_profile(p, (32, 32)))


def test_exact_group_risk_is_selected_for_its_options():
    # the variance (HT) or variance / risk (Wiener) model in the windowed-
    # synthesis domain with group scope and k > 0 takes the exact
    # covariance-aware group risk (bmndalgo.py:359-377, 752-758)
    from bmnd_amd.engine import check_supported, exact_risk_stage
    from bmnd_amd.profilevalidation import resolve_profile
    p = _poisson()
    assert not exact_risk_stage(p, "ht", 4) and not exact_risk_stage(p, "wiener", 4)
    p.wiener_weight_domain, p.wiener_weight_scope = "windowed_synthesis", "group"
    p.ht_weight_domain, p.ht_weight_scope = "windowed_synthesis", "group"
    assert exact_risk_stage(p, "ht", 4) and exact_risk_stage(p, "wiener", 4)
    assert not exact_risk_stage(p, "wiener", 0)
    p.wiener_weight_model = "risk"
    assert exact_risk_stage(p, "wiener", 4)
    p.ht_weight_model = "classic"
    assert not exact_risk_stage(p, "ht", 4)
    check_supported(p, resolve_profile(p, (32, 32)))


def test_gaussian_profile_support():
    # the Gaussian model runs with SSD matching; a Poisson distance, or the
    # exact covariance-aware group risk, is refused before any device work
    from bmnd_amd.engine import check_supported
    from bmnd_amd.profilevalidation import resolve_profile
    prof = BM3DProfile()
    check_supported(prof, resolve_profile(prof, (32, 32)))
    prof.blockmatch_distance = "poisson_deviance"
    with pytest.raises(NotImplementedError, match="'ssd'"):
        check_supported(prof, resolve_profile(prof, (32, 32)))
    prof = BM3DProfile()
    prof.ht_weight_domain, prof.ht_weight_scope = "windowed_synthesis", "group"
    with pytest.raises(NotImplementedError, match="exact covariance-aware"):
        check_supported(prof, resolve_profile(prof, (32, 32)))
    # the conventional nf path has k = 0, where that risk is the diagonal one
    prof.nf = 0
    check_supported(prof, resolve_profile(prof, (32, 32)))


def test_exact_std_work_counts_union_voxels():
    """The exact-variance work counter (used for the roofline) against a
    brute-force union of the member patches."""
    import torch
    from bmnd_amd.engine import exact_std_work, make_plan
    plan = make_plan(BM4DProfile(noise_model=NoiseModel.POISSON), "wiener", 3)
    shape3 = (14, 15, 16)
    b = plan.block3
    c = [shape3[d] - b[d] + 1 for d in range(3)]
    rng = np.random.default_rng(1)
    R = 12
    cnt = rng.integers(1, plan.max_m + 1, R)
    sel = np.full((R, plan.max_m), -1, np.int64)
    for r in range(R):
        sel[r, : cnt[r]] = rng.choice(int(np.prod(c)), cnt[r], replace=False)
    w = exact_std_work(plan, shape3, torch.from_numpy(sel.astype(np.int32).ravel()),
                       torch.from_numpy(cnt.astype(np.int32)), 4, chunk=5)
    P = int(np.prod(b))
    flops = union = 0
    for r in range(R):
        vox = set()
        for lin in sel[r, : cnt[r]]:
            x2, x1, x0 = lin % c[2], (lin // c[2]) % c[1], lin // (c[2] * c[1])
            for p in range(P):
                p2, p1, p0 = p % b[2], (p // b[2]) % b[1], p // (b[2] * b[1])
                vox.add(((x0 + p0) * shape3[1] + x1 + p1) * shape3[2] + x2 + p2)
        union += len(vox)
        flops += min(4, cnt[r]) * P * (2 * cnt[r] * P + 3 * len(vox))
    assert w["entries"] == int(cnt.sum()) * P
    assert w["union_voxels"] == union
    assert w["flops"] == flops
