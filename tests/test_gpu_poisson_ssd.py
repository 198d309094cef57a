"""The SSD distance under the Poisson model on the GPU against the CPU
oracle, bit for bit: the stage-1 matcher with the Poisson gamma correction
(blockmatching.py:432-443, 1140-1145) over every patch origin of small
inputs of the three compiled geometries, integer and non-integer counts,
several gammas and count scales; then both stages of bmnd() with
blockmatch_distance "ssd" (bmndalgo.py:285-322, 721-741): the match lists,
the pilot and the output."""

import ctypes

import numpy as np
import pytest
import torch

import oracle
from bmnd_amd import BM3DProfile, BM4DProfile, NoiseModel, _lib, phantoms
from bmnd_amd.engine import PoissonPipeline

from test_gpu_parity import _compare

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)


def _input(shape, seed, scale, jitter):
    rng = np.random.default_rng(seed)
    lam = rng.uniform(0.5, 12.0, size=shape)
    vol = (rng.poisson(lam) * scale).astype(np.float32)
    if jitter:
        vol = vol + rng.uniform(-0.3, 0.3, vol.shape).astype(np.float32)
    return vol


GEOMS = {  # (shape, block, radius, max stack)
    "image": ((40, 46), (8, 8), (19, 19), 16),
    "vol4": ((18, 19, 20), (4, 4, 4), (7, 7, 7), 16),
    "vol5": ((19, 18, 20), (5, 5, 5), (7, 7, 7), 32),
}


@pytest.mark.parametrize("geom", sorted(GEOMS))
@pytest.mark.parametrize("gamma,scale,jitter", [(3.0, 1.0, False), (0.5, 0.37, True),
                                                (1.25, 2.5, True)])
def test_stage1_matcher_matches_oracle(geom, gamma, scale, jitter):
    shape, block, radius, max_m = GEOMS[geom]
    print(f"poisson ssd matcher: {geom} gamma {gamma} scale {scale} jitter {jitter}", flush=True)
    vol = _input(shape, 11 + len(shape), scale, jitter)
    nd = vol.ndim
    counts = tuple(shape[d] - block[d] + 1 for d in range(nd))
    refs = np.stack(np.unravel_index(np.arange(int(np.prod(counts))), counts), axis=1)
    thr = float(np.float32(0.05 * float(vol.max() - vol.min()) ** 2 * np.prod(block) / 16))
    o_sel, o_cnt, o_acc = oracle.blockmatch_ssd_poisson(vol, block, radius, refs, thr, max_m, 2,
                                                        gamma, count_scale=scale)
    pad = 3 - nd
    shape3 = (ctypes.c_int64 * 3)(*((1,) * pad + shape))
    block3 = (ctypes.c_int32 * 3)(*((1,) * pad + block))
    radius3 = (ctypes.c_int32 * 3)(*((0,) * pad + radius))
    R = len(refs)
    lin = torch.arange(R, dtype=torch.int32, device=DEV)
    d_vol = torch.from_numpy(vol).to(DEV)
    sel = torch.empty(R * max_m, dtype=torch.int32, device=DEV)
    cnt = torch.empty(R, dtype=torch.int32, device=DEV)
    acc = torch.empty(R, dtype=torch.int32, device=DEV)
    _lib.check(_lib.lib().bmnd_hip_blockmatch_ssd_poisson(
        d_vol.data_ptr(), shape3, block3, radius3, lin.data_ptr(), R, thr, max_m, 2, 1, gamma,
        scale, sel.data_ptr(), cnt.data_ptr(), acc.data_ptr(), None))
    torch.cuda.synchronize()
    g_sel = sel.view(R, max_m).cpu().numpy()
    g_cnt, g_acc = cnt.cpu().numpy(), acc.cpu().numpy()
    assert np.array_equal(g_cnt, o_cnt)
    assert np.array_equal(g_acc, o_acc.astype(np.int32))
    differing = sum(1 for r in range(R) if not np.array_equal(g_sel[r, :g_cnt[r]],
                                                               o_sel[r, :o_cnt[r]]))
    print(f"  {R} references, accepted mean {o_acc.mean():.1f}, {differing} differing",
          flush=True)
    assert differing == 0


@pytest.mark.parametrize("kind,gamma,scale", [("image", 3.0, 1.0), ("volume", 3.0, 1.0),
                                              ("image", 0.0, 1.0), ("volume", 1.5, 0.5)])
def test_bmnd_poisson_ssd_matches_oracle(kind, gamma, scale):
    print(f"poisson ssd bmnd: {kind} gamma {gamma} scale {scale}", flush=True)
    if kind == "volume":
        _, noisy = phantoms.ellipsoid_volume((24, 26, 22), seed=81, peak=10.0, n_ellipsoids=6,
                                             axes_range=(3.0, 10.0))
        prof = BM4DProfile(noise_model=NoiseModel.POISSON)
    else:
        _, noisy = phantoms.blob_image(64, seed=82, peak=20.0)
        prof = BM3DProfile(noise_model=NoiseModel.POISSON)
    noisy = (noisy * np.float32(scale)).astype(np.float32)
    prof.blockmatch_distance = "ssd"
    prof.gamma = gamma
    prof.poisson_count_scale = scale
    params = oracle.default_params(noisy.ndim, noisy.shape)
    params.count_scale = scale
    oracle.set_poisson_ssd(params, ht_match=prof.ht_match_threshold,
                           wi_match=prof.wiener_match_threshold, gamma=gamma)
    stats, out, o_out = _compare(noisy, prof, params, "auto-fast")
    assert stats["modes"] == {"ht": 11 if gamma else 10, "wiener": 10}
    assert stats["ht_groups_differing"] == 0
    assert stats["wiener_groups_differing"] == 0
    assert stats["pilot_bits_differing"] == 0
    assert stats["out_bits_differing"] == 0
