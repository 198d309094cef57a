"""The stationary Gaussian path on the GPU against the CPU oracle with the
same noise description (host tables), bit for bit: both stages' match lists
(SSD with the stage-1 noise-bias correction), the exact-plane group
variances, the pilot, the output and the sigma maps.  Cases: the profiles'
nf domains with k = 4 and gamma = 3, the conventional nf path (k = 0, no
correction), other k, an nf domain shorter than the Wiener block (several
patch entries per domain cell), a coloured sigma_psd, the estimated sigma,
soft thresholding with the risk and classic weight models
(bmndalgo.py:203-1063, variance.py:452-553, 1791-2106)."""

import ctypes

import numpy as np
import pytest
import torch

import oracle
from bmnd_amd import bmnd
from bmnd_amd import _lib
from bmnd_amd.enums import AggregationWeightDomain, AggregationWeightModel, AggregationWeightScope
from bmnd_amd.gaussian import GaussianPipeline
from bmnd_amd.transforms import group_matrices

from gaussian_helpers import noisy_input, oracle_params

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)


def _coloured_psd(shape, seed=0):
    f = np.meshgrid(*[np.fft.fftfreq(s) for s in shape], indexing="ij")
    psd = np.ones(shape, np.float64)
    for a, g in enumerate(f):
        psd += (2.0 + a) * np.abs(g)
    return (psd * np.prod(shape) * 0.01).astype(np.float32)


def _lists_differ(gs, gc, os_, oc):
    return sum(1 for r in range(gs.shape[0])
               if gc[r] != oc[r] or not np.array_equal(gs[r, :gc[r]], os_[r, :oc[r]]))


def _run_both(noisy, prof, sigma=0.1, sigma_psd=None, sigma_maps=False):
    pipe = GaussianPipeline(prof, noisy.shape, DEV)
    ht, wi, g = pipe.run(torch.from_numpy(noisy).to(DEV), sigma=sigma, sigma_psd=sigma_psd,
                         sigma_maps=sigma_maps, vol_host=noisy)
    torch.cuda.synchronize()
    tables = pipe.tables(torch.from_numpy(noisy).to(DEV), sigma, sigma_psd)
    params = oracle_params(prof, noisy.shape, tables.model)
    if prof.ht_use_soft_thresholding:
        params.ht_soft = 1
    for st in ("ht", "wiener"):
        oracle.set_weight_options(params, st, getattr(prof, f"{st}_weight_model").value,
                                  getattr(prof, f"{st}_weight_domain")
                                  is AggregationWeightDomain.WINDOWED_SYNTHESIS,
                                  getattr(prof, f"{st}_weight_scope") is AggregationWeightScope.GROUP)
    o_out, o_pilot, o_hmap, o_wmap, (o_hs, o_hc), (o_ws, o_wc) = oracle.bmnd_gaussian_sigma(
        noisy, params, capture_matches=True)
    stats = {
        "ht_groups_differing": _lists_differ(ht.sel.cpu().numpy(), ht.counts.cpu().numpy(),
                                             o_hs, o_hc),
        "wiener_groups_differing": _lists_differ(wi.sel.cpu().numpy(), wi.counts.cpu().numpy(),
                                                 o_ws, o_wc),
        "pilot_bits_differing": int(np.sum(ht.est.cpu().numpy().view(np.int32)
                                           != o_pilot.view(np.int32))),
        "out_bits_differing": int(np.sum(wi.est.cpu().numpy().view(np.int32)
                                         != o_out.view(np.int32))),
        "k": pipe.k, "gamma": pipe.gamma, "domain": tables.model.variance_domain_shape,
    }
    if sigma_maps:
        stats["ht_map_bits_differing"] = int(np.sum(ht.sigma_map.cpu().numpy().view(np.int32)
                                                    != o_hmap.view(np.int32)))
        stats["wiener_map_bits_differing"] = int(np.sum(wi.sigma_map.cpu().numpy().view(np.int32)
                                                        != o_wmap.view(np.int32)))
        stats["global"] = (g, tables.model.global_sigma)
    print(stats, flush=True)
    return stats, wi.est.cpu().numpy(), o_out


def _assert_identical(stats):
    for key, v in stats.items():
        if key.endswith("_differing"):
            assert v == 0, (key, stats)


@pytest.mark.parametrize("kind", ["volume", "image"])
@pytest.mark.parametrize("case", ["default", "conventional", "k1", "k2", "short-nf", "coloured",
                                  "estimated", "soft-risk", "classic-windowed"])
def test_pipeline_matches_oracle(kind, case):
    print(f"gaussian pipeline: {kind} {case}", flush=True)
    truth, noisy, prof = noisy_input(kind, 0.1, seed=3)
    sigma, psd = 0.1, None
    if case == "conventional":
        prof.nf = 0
    elif case in ("k1", "k2"):
        prof.k = int(case[1])
    elif case == "short-nf":
        prof.nf = (4,) * noisy.ndim
    elif case == "coloured":
        sigma, psd = None, _coloured_psd(noisy.shape)
    elif case == "estimated":
        sigma = None
    elif case == "soft-risk":
        prof.ht_use_soft_thresholding = True
        prof.wiener_weight_model = AggregationWeightModel.RISK
        prof.ht_weight_scope = AggregationWeightScope.GROUP
    elif case == "classic-windowed":
        prof.ht_weight_model = AggregationWeightModel.CLASSIC
        prof.wiener_weight_domain = AggregationWeightDomain.WINDOWED_SYNTHESIS
    stats, out, _ = _run_both(noisy, prof, sigma, psd)
    _assert_identical(stats)
    if case in ("default", "conventional"):
        assert np.mean((out - truth) ** 2) < 0.3 * np.mean((noisy - truth) ** 2)


@pytest.mark.parametrize("kind", ["volume", "image"])
def test_sigma_maps_match_oracle(kind):
    print(f"gaussian sigma maps: {kind}", flush=True)
    _, noisy, prof = noisy_input(kind, 0.2, seed=4)
    stats, _, _ = _run_both(noisy, prof, 0.2, None, sigma_maps=True)
    _assert_identical(stats)
    assert stats["global"][0] == stats["global"][1]


@pytest.mark.parametrize("kind", ["volume", "image"])
def test_group_variance_matches_oracle_per_group(kind):
    # the device's exact planes against the oracle's per-group function on
    # the HT and Wiener groups of a run
    print(f"gaussian group variance: {kind}", flush=True)
    _, noisy, prof = noisy_input(kind, 0.1, seed=5)
    pipe = GaussianPipeline(prof, noisy.shape, DEV)
    vol = torch.from_numpy(noisy).to(DEV)
    tables = pipe.tables(vol, 0.1)
    L = _lib.lib()
    for name in ("ht", "wiener"):
        out = pipe.stage(name, vol, vol, tables, pilot=vol if name == "wiener" else None)
        plan = pipe.plans[name]
        R = out.counts.numel()
        kcap = min(pipe.k, plan.max_m)
        exact = torch.empty(R * kcap * plan.P, dtype=torch.float32, device=DEV)
        mats, _ = plan.device_constants(DEV)
        s = ctypes.c_void_p(torch.cuda.current_stream(DEV).cuda_stream)
        _lib.check(L.bmnd_hip_gaussian_variance(
            ctypes.c_void_p(tables.cov[name].data_ptr()), (ctypes.c_int32 * 3)(*tables.domain3),
            (ctypes.c_int64 * 3)(*pipe.shape3), (ctypes.c_int32 * 3)(*plan.block3),
            (ctypes.c_int32 * 3)(*plan.radius3), ctypes.c_void_p(out.sel.data_ptr()),
            ctypes.c_void_p(out.counts.data_ptr()), R, plan.max_m,
            ctypes.c_void_p(mats.data_ptr()), pipe.k, ctypes.c_void_p(exact.data_ptr()), s))
        ex = exact.cpu().numpy().reshape(R, kcap, plan.P)
        sel, cnt = out.sel.cpu().numpy(), out.counts.cpu().numpy()
        maps = tables.cov[name].cpu().numpy().T.reshape((plan.P,) + tables.model.variance_domain_shape)
        counts = [noisy.shape[d] - plan.block[d] + 1 for d in range(noisy.ndim)]
        rng = np.random.default_rng(0)
        diff = 0
        picks = rng.choice(R, size=min(R, 150), replace=False)
        for r in picks:
            n = int(cnt[r])
            pos = np.stack(np.unravel_index(sel[r, :n], counts), axis=1)
            want = oracle.gaussian_group_variance(maps, pos, group_matrices(n)[0], pipe.k)
            q = min(pipe.k, n)
            diff += int(np.sum(ex[r, :q].view(np.int32) != want[:q].view(np.int32)))
        print(f"  {name}: {len(picks)} groups, {diff} values differing", flush=True)
        assert diff == 0


def test_bmnd_routes_gaussian_profiles_and_callbacks():
    _, noisy, prof = noisy_input("image", 0.1, seed=6)
    events = []
    out = bmnd(noisy, prof, sigma=0.1,
               stage_callback=lambda st, est: events.append(("stage", st, est.shape)),
               blockmatch_callback=lambda st, res: events.append(("match", st, len(res.groups))))
    assert [e[:2] for e in events] == [("match", "ht"), ("stage", "ht"), ("match", "wiener"),
                                       ("stage", "wiener")]
    pipe = GaussianPipeline(prof, noisy.shape, DEV)
    _, wi, _ = pipe.run(torch.from_numpy(noisy).to(DEV), sigma=0.1)
    assert np.array_equal(out.view(np.int32), wi.est.cpu().numpy().view(np.int32))
    out2, maps = bmnd(noisy, prof, sigma=0.1, return_sigma_map=True)
    assert np.array_equal(out2.view(np.int32), out.view(np.int32))
    assert set(maps) == {"global", "ht", "wiener"}
    # sigma_psd given as a scalar is that scalar, read as float32, taken as
    # sigma (psd.py:33-55, 115-131)
    out3 = bmnd(noisy, prof, sigma_psd=0.1)
    out4 = bmnd(noisy, prof, sigma=float(np.float32(0.1)))
    assert np.array_equal(out3.view(np.int32), out4.view(np.int32))


@pytest.mark.parametrize("kind,shape", [("image", (256, 256)), ("volume", (48, 44, 40))])
def test_larger_inputs_match_oracle(kind, shape):
    print(f"gaussian larger: {kind} {shape}", flush=True)
    truth, noisy, prof = noisy_input(kind, 0.15, seed=8, shape=shape)
    stats, out, _ = _run_both(noisy, prof, 0.15)
    _assert_identical(stats)
    assert np.mean((out - truth) ** 2) < 0.3 * np.mean((noisy - truth) ** 2)


def test_batch_runs_each_input_as_bmnd():
    from bmnd_amd import bmnd_batch
    inputs = [noisy_input("image", 0.1, seed=s, shape=(n, n))[1] for s, n in ((9, 48), (10, 72))]
    prof = noisy_input("image", 0.1, seed=9, shape=(48, 48))[2]
    outs = bmnd_batch(inputs + [np.zeros((40, 40), np.float32)], prof, sigma=0.1)
    for x, o in zip(inputs, outs):
        assert np.array_equal(o.view(np.int32), bmnd(x, prof, sigma=0.1).view(np.int32))
    assert not np.any(outs[2])


def _estimates(vol):
    from bmnd_amd.psd import estimate_gaussian_global_sigma
    pipe = GaussianPipeline(noisy_input("volume", 0.1)[2], (16, 16, 16), DEV)
    dev = pipe.estimate_sigma(torch.from_numpy(np.ascontiguousarray(vol, np.float32)).to(DEV))
    return dev, estimate_gaussian_global_sigma(vol)


@pytest.mark.parametrize("case", ["odd", "even", "zeros", "ties", "constant", "one", "negative"])
def test_device_sigma_estimate_equals_numpy(case):
    # the device median / MAD against numpy's on the host
    # (sigmaestimation.py:7-45), exactly
    rng = np.random.default_rng(11)
    if case == "odd":
        vol = rng.normal(0.3, 0.2, size=(31, 33, 35)).astype(np.float32)
    elif case == "even":
        vol = rng.normal(0.3, 0.2, size=(32, 33, 34)).astype(np.float32)
    elif case == "zeros":
        vol = rng.normal(0.0, 1.0, size=(40, 40)).astype(np.float32)
        vol[rng.random(vol.shape) < 0.4] = 0.0
    elif case == "ties":
        vol = rng.integers(-3, 5, size=(64, 64)).astype(np.float32)
    elif case == "constant":
        vol = np.full((20, 20, 20), 2.5, np.float32)  # MAD 0: the fallbacks
    elif case == "one":
        vol = np.zeros((16, 16), np.float32)
        vol[3, 4] = 7.0
    else:
        vol = -np.abs(rng.normal(size=(50, 51))).astype(np.float32)
    dev, host = _estimates(vol)
    print(f"sigma estimate {case}: device {dev!r} numpy {host!r}", flush=True)
    assert dev == host
