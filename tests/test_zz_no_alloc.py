"""The no-allocation stage entry point (bmnd_hip_stage_noalloc, scratch in a
static device array) against the workspace stage driver (bmnd_hip_stage):
identical match lists, counts, accepted counts and estimates for both
stages, on two calls in a row with the second input smaller than the first
(the scratch keeps the first call's contents), on a volume and an image;
and a stage larger than the scratch is refused with a capacity error."""

import ctypes

import numpy as np
import pytest
import torch

from bmnd_amd import BM3DProfile, BM4DProfile, NoiseModel, _lib, phantoms
from bmnd_amd.engine import PoissonPipeline, _ptr

pytestmark = pytest.mark.gpu

ECAPACITY = -3


def _run(pipe, name, match_vol, noisy, var, noalloc):
    L = _lib.lib()
    plan = pipe.plans[name]
    shape = (ctypes.c_int64 * 3)(*pipe.shape3)
    block = (ctypes.c_int32 * 3)(*plan.block3)
    radius = (ctypes.c_int32 * 3)(*plan.radius3)
    refs = pipe._ref_lin(plan)
    R = refs.numel()
    dev = pipe.device
    est = torch.full_like(noisy, float("nan"))
    sel = torch.full((R * plan.max_m,), -7, dtype=torch.int32, device=dev)
    cnt = torch.full((R,), -7, dtype=torch.int32, device=dev)
    acc = torch.full((R,), -7, dtype=torch.int32, device=dev)
    mats, kai = plan.device_constants(dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    stage_arg = int(name == "wiener")
    common = (stage_arg, _ptr(match_vol), _ptr(noisy), _ptr(var), _ptr(noisy), shape, block,
              radius, _ptr(refs), R, plan.threshold, plan.max_m, plan.min_m, pipe.count_scale,
              -1, _ptr(mats), _ptr(kai), plan.shrink, pipe.k)
    outs = (_ptr(est), _ptr(sel), _ptr(cnt), _ptr(acc), stream)
    if noalloc:
        _lib.check(L.bmnd_hip_stage_noalloc(*common, *outs))
    else:
        need = L.bmnd_hip_stage_workspace(shape, block, R, plan.max_m, pipe.k)
        ws = torch.empty(need, dtype=torch.uint8, device=dev)
        _lib.check(L.bmnd_hip_stage(*common, _ptr(ws), need, *outs))
    torch.cuda.synchronize()
    return est, sel, cnt, acc


def _both_stages_equal(prof, noisy):
    dev = torch.device("cuda", 0)
    pipe = PoissonPipeline(prof, noisy.shape, dev)
    v = torch.from_numpy(np.ascontiguousarray(noisy)).to(dev)
    var1 = pipe.variance(v)
    ref = _run(pipe, "ht", v, v, var1, noalloc=False)
    got = _run(pipe, "ht", v, v, var1, noalloc=True)
    for a, b in zip(ref, got):
        assert torch.equal(a.view(torch.int32), b.view(torch.int32))
    pilot = ref[0]
    var2 = pipe.variance(pilot)
    ref = _run(pipe, "wiener", pilot, v, var2, noalloc=False)
    got = _run(pipe, "wiener", pilot, v, var2, noalloc=True)
    for a, b in zip(ref, got):
        assert torch.equal(a.view(torch.int32), b.view(torch.int32))
    # the Wiener stage matched on the (non-integer) pilot: the fast matcher,
    # whose patch sums are in the scratch
    assert not torch.equal(pilot, torch.round(pilot))


@pytest.mark.parametrize("kind", ["volume", "image"])
def test_noalloc_stage_equals_workspace_stage_twice(kind):
    print(f"no-alloc stage: {kind}", flush=True)
    _lib.check(_lib.lib().bmnd_hip_init())
    if kind == "volume":
        prof = BM4DProfile(noise_model=NoiseModel.POISSON)
        big = phantoms.ellipsoid_volume((26, 28, 24), seed=71, peak=10.0, n_ellipsoids=6,
                                        axes_range=(3.0, 10.0))[1]
        small = phantoms.ellipsoid_volume((16, 18, 14), seed=72, peak=4.0, n_ellipsoids=4,
                                          axes_range=(2.0, 6.0))[1]
    else:
        prof = BM3DProfile(noise_model=NoiseModel.POISSON)
        big = phantoms.blob_image(88, seed=73, peak=20.0)[1]
        small = phantoms.square_image(3)[1]
    _both_stages_equal(prof, big)
    _both_stages_equal(prof, small)


def test_noalloc_stage_refuses_what_the_scratch_cannot_hold():
    print("no-alloc stage: capacity", flush=True)
    L = _lib.lib()
    _lib.check(L.bmnd_hip_init())
    dummy = ctypes.c_void_p(16)
    shape = (ctypes.c_int64 * 3)(128, 128, 128)
    block = (ctypes.c_int32 * 3)(5, 5, 5)
    radius = (ctypes.c_int32 * 3)(7, 7, 7)
    rc = L.bmnd_hip_stage_noalloc(1, dummy, dummy, dummy, dummy, shape, block, radius, dummy,
                                  200000, 0.7689, 32, 2, 1.0, -1, dummy, dummy, 0.4, 4, dummy,
                                  dummy, dummy, dummy, None)
    assert rc == ECAPACITY
    assert "bytes of scratch" in L.bmnd_hip_last_error().decode()
