"""The fused stage driver of the C-ABI (bmnd_hip_stage with a caller-owned
workspace) against the engine's call-by-call stage: identical match lists,
counts, accepted counts and estimates for both stages of a volume and an
image, with the automatic matching mode, and with the Pearson distance
(matching mode 5) and the classic Wiener gain (stage argument 2)."""

import ctypes

import numpy as np
import pytest
import torch

from bmnd_amd import BM3DProfile, BM4DProfile, NoiseModel, _lib, phantoms
from bmnd_amd.engine import PoissonPipeline, _ptr

pytestmark = pytest.mark.gpu


def _driver_stage(pipe, name, match_vol, noisy, var, fallback, mode=-1, stage_arg=None):
    L = _lib.lib()
    plan = pipe.plans[name]
    shape = (ctypes.c_int64 * 3)(*pipe.shape3)
    block = (ctypes.c_int32 * 3)(*plan.block3)
    radius = (ctypes.c_int32 * 3)(*plan.radius3)
    refs = pipe._ref_lin(plan)
    R = refs.numel()
    need = L.bmnd_hip_stage_workspace(shape, block, R, plan.max_m, pipe.k)
    assert need > 0
    dev = pipe.device
    ws = torch.empty(need, dtype=torch.uint8, device=dev)
    est = torch.empty_like(noisy)
    sel = torch.empty(R * plan.max_m, dtype=torch.int32, device=dev)
    cnt = torch.empty(R, dtype=torch.int32, device=dev)
    acc = torch.empty(R, dtype=torch.int32, device=dev)
    mats, kai = plan.device_constants(dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    if stage_arg is None:
        stage_arg = int(name == "wiener")
    _lib.check(L.bmnd_hip_stage(stage_arg, _ptr(match_vol), _ptr(noisy), _ptr(var),
                                _ptr(fallback), shape, block, radius, _ptr(refs), R,
                                plan.threshold, plan.max_m, plan.min_m, pipe.count_scale, mode,
                                _ptr(mats), _ptr(kai), plan.shrink, pipe.k, _ptr(ws), need,
                                _ptr(est), _ptr(sel), _ptr(cnt), _ptr(acc), stream))
    return est, sel.view(R, plan.max_m), cnt, acc


@pytest.mark.parametrize("kind", ["volume", "image"])
def test_stage_driver_matches_engine(kind):
    print(f"stage driver: {kind}", flush=True)
    if kind == "volume":
        _, noisy = phantoms.ellipsoid_volume((40, 44, 36), seed=61, peak=10.0, n_ellipsoids=8,
                                             axes_range=(3.0, 12.0))
        prof = BM4DProfile(noise_model=NoiseModel.POISSON)
    else:
        _, noisy = phantoms.blob_image(96, seed=62, peak=20.0)
        prof = BM3DProfile(noise_model=NoiseModel.POISSON)
    dev = torch.device("cuda", 0)
    pipe = PoissonPipeline(prof, noisy.shape, dev)
    v = torch.from_numpy(np.ascontiguousarray(noisy)).to(dev)
    ht, wi = pipe.run(v)
    var1 = pipe.variance(v)
    d_est, d_sel, d_cnt, d_acc = _driver_stage(pipe, "ht", v, v, var1, v)
    assert torch.equal(d_sel, ht.sel) and torch.equal(d_cnt, ht.counts)
    assert torch.equal(d_acc, ht.accepted)
    assert torch.equal(d_est.view(torch.int32), ht.est.view(torch.int32))
    var2 = pipe.variance(ht.est)
    w_est, w_sel, w_cnt, _ = _driver_stage(pipe, "wiener", ht.est, v, var2, v)
    assert torch.equal(w_sel, wi.sel) and torch.equal(w_cnt, wi.counts)
    assert torch.equal(w_est.view(torch.int32), wi.est.view(torch.int32))


def test_stage_driver_with_options():
    print("stage driver: pearson distance, classic gain", flush=True)
    _, noisy = phantoms.blob_image(80, seed=63, peak=15.0)
    prof = BM3DProfile(noise_model=NoiseModel.POISSON)
    prof.blockmatch_distance = "pearson"
    prof.poisson_wiener_gain_mode = "classic"
    dev = torch.device("cuda", 0)
    pipe = PoissonPipeline(prof, noisy.shape, dev)
    assert pipe.wiener_arg == 2
    v = torch.from_numpy(np.ascontiguousarray(noisy)).to(dev)
    ht, wi = pipe.run(v)
    var1 = pipe.variance(v)
    d_est, d_sel, d_cnt, d_acc = _driver_stage(pipe, "ht", v, v, var1, v, mode=5, stage_arg=0)
    assert torch.equal(d_sel, ht.sel) and torch.equal(d_cnt, ht.counts)
    assert torch.equal(d_acc, ht.accepted)
    assert torch.equal(d_est.view(torch.int32), ht.est.view(torch.int32))
    var2 = pipe.variance(ht.est)
    w_est, w_sel, w_cnt, _ = _driver_stage(pipe, "wiener", ht.est, v, var2, v, mode=5,
                                           stage_arg=pipe.wiener_arg)
    assert torch.equal(w_sel, wi.sel) and torch.equal(w_cnt, wi.counts)
    assert torch.equal(w_est.view(torch.int32), wi.est.view(torch.int32))


def test_stage_driver_explicit_count_table_mode():
    """Mode 2 given explicitly: the driver reads the count range itself and
    matches like the automatic choice; a non-integer input is refused."""
    print("stage driver: explicit count-table mode", flush=True)
    _, noisy = phantoms.ellipsoid_volume((30, 28, 26), seed=64, peak=10.0, n_ellipsoids=6,
                                         axes_range=(3.0, 10.0))
    prof = BM4DProfile(noise_model=NoiseModel.POISSON)
    dev = torch.device("cuda", 0)
    pipe = PoissonPipeline(prof, noisy.shape, dev)
    v = torch.from_numpy(np.ascontiguousarray(noisy)).to(dev)
    ht, _ = pipe.run(v)
    assert pipe.last_modes["ht"] == 2
    var1 = pipe.variance(v)
    d_est, d_sel, d_cnt, d_acc = _driver_stage(pipe, "ht", v, v, var1, v, mode=2, stage_arg=0)
    assert torch.equal(d_sel, ht.sel) and torch.equal(d_cnt, ht.counts)
    assert torch.equal(d_acc, ht.accepted)
    assert torch.equal(d_est.view(torch.int32), ht.est.view(torch.int32))
    with pytest.raises(_lib.HipError, match="integer counts <= 100"):
        _driver_stage(pipe, "ht", v + 0.5, v, var1, v, mode=2, stage_arg=0)
