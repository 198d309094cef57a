"""GPU sigma maps (bmnd(..., return_sigma_map=True)) against the CPU oracle,
bit for bit: the global sigma, the HT map (per-coefficient limited std
scattered with the aggregation weights) and the Wiener map (per-group numpy
mean of the stds), with the estimate unchanged by asking for the maps.  The
global-sigma kernel is also checked at the full 256^3 size against numpy's
own np.mean (a size-independent check: numpy is the arithmetic the reference
runs), and on the zero-floor branches (positive subset, sparse trimmed
mean) against the oracle."""

import ctypes

import numpy as np
import pytest
import torch

import oracle
from bmnd_amd import BM3DProfile, BM4DProfile, NoiseModel, _lib, bmnd, phantoms

pytestmark = pytest.mark.gpu


def _same_bits(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.int32),
                          np.asarray(b, np.float32).view(np.int32))


def _check_api(noisy, prof, ndim):
    print(f"sigma maps: shape {noisy.shape}", flush=True)
    plain = bmnd(noisy, prof)
    out, maps = bmnd(noisy, prof, return_sigma_map=True)
    o_out, o_maps = oracle.bmnd_poisson_sigma_maps(noisy, oracle.default_params(ndim, noisy.shape))
    assert _same_bits(out, plain)
    assert _same_bits(out, o_out)
    assert set(maps) == {"global", "ht", "wiener"}
    assert isinstance(maps["global"], float)
    assert maps["global"] == o_maps["global"]
    for stage in ("ht", "wiener"):
        assert maps[stage].shape == noisy.shape and maps[stage].dtype == np.float32
        diff = int(np.sum(maps[stage].view(np.int32) != o_maps[stage].view(np.int32)))
        assert diff == 0, f"{stage}: {diff} voxels differ"


def test_sigma_maps_bm4d_volume():
    _, noisy = phantoms.ellipsoid_volume((24, 26, 22), seed=31, peak=10.0, n_ellipsoids=5,
                                         axes_range=(3.0, 10.0))
    _check_api(noisy, BM4DProfile(noise_model=NoiseModel.POISSON), 3)


def test_sigma_maps_bm3d_image():
    _, noisy = phantoms.blob_image(53, seed=32, peak=10.0)
    noisy = np.ascontiguousarray(noisy[:41, :53])
    _check_api(noisy, BM3DProfile(noise_model=NoiseModel.POISSON), 2)


def test_sigma_maps_reference_stability_image():
    """tests/test_bmnd_robustness.py:49-67 asks for finite maps at rate 5 on
    an 8x8 image."""
    noisy = np.random.default_rng(0).poisson(5.0, size=(8, 8)).astype(np.float32)
    _check_api(noisy, BM3DProfile(noise_model=NoiseModel.POISSON), 2)


def _global_sigma(var: np.ndarray, floor: float) -> float:
    L = _lib.lib()
    t = torch.from_numpy(np.ascontiguousarray(var, np.float32).ravel()).cuda()
    g = ctypes.c_double(0.0)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(L.bmnd_hip_poisson_global_sigma(ctypes.c_void_p(t.data_ptr()), t.numel(), floor,
                                               ctypes.byref(g), stream))
    return g.value


def test_global_sigma_full_size_matches_numpy():
    print("global sigma: 256^3", flush=True)
    rng = np.random.default_rng(33)
    var = np.maximum(rng.poisson(10.0, (256, 256, 256)).astype(np.float32), np.float32(1e-10))
    want = float(np.sqrt(max(float(np.mean(var, dtype=np.float32)), 1e-10)))
    assert _global_sigma(var, 1e-10) == want


@pytest.mark.parametrize("rate", [0.7, 0.03])
def test_global_sigma_zero_floor_branches(rate):
    """Zero floor: the positive subset (rate 0.7) and the sparse trimmed mean
    of sorted values (rate 0.03, under a tenth positive)."""
    print(f"global sigma: zero floor, rate {rate}", flush=True)
    rng = np.random.default_rng(34)
    var = rng.poisson(rate, (96, 80, 70)).astype(np.float32)
    var[var > 0] *= rng.gamma(2.0, 1.0, int(np.sum(var > 0))).astype(np.float32)
    assert _global_sigma(var, 0.0) == oracle.poisson_global_sigma(var, 0.0)


def test_global_sigma_small_and_degenerate():
    for n in (1, 7, 129, 8191, 8193):
        var = (np.random.default_rng(n).random(n) + 0.5).astype(np.float32)
        assert _global_sigma(var, 1e-10) == oracle.poisson_global_sigma(var, 1e-10)
    assert _global_sigma(np.zeros(5000, np.float32), 0.0) == np.sqrt(1e-10)
