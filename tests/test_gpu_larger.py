"""GPU parity on larger inputs than tests/test_gpu_parity.py uses.

A 160x160 low-count image makes the fast matcher's certified evaluation
lists long (hundreds of candidates per reference, so each candidate gets one
or two lanes), and a 48^3 volume gives unclipped BM4D search boxes for most
references; both are compared with the CPU oracle bit for bit.  At 128^3,
beyond what the oracle finishes quickly, the checks are size-independent:
two runs are bit-identical, and a batch of one input equals the single-input
pipeline."""

import numpy as np
import pytest
import torch

import oracle
from bmnd_amd import BM3DProfile, BM4DProfile, NoiseModel, phantoms
from bmnd_amd.engine import BatchPipeline, PoissonPipeline

from test_gpu_parity import _compare

pytestmark = pytest.mark.gpu


def _assert_identical(stats):
    assert stats["ht_groups_differing"] == 0
    assert stats["wiener_groups_differing"] == 0
    assert stats["pilot_bits_differing"] == 0
    assert stats["out_bits_differing"] == 0


@pytest.mark.parametrize("peak", [1.0, 5.0])
def test_low_count_image_with_long_exact_lists(peak):
    print(f"160x160 image, peak {peak}", flush=True)
    _, noisy = phantoms.blob_image(160, seed=51, peak=peak)
    prof = BM3DProfile(noise_model=NoiseModel.POISSON)
    stats, _, _ = _compare(noisy, prof, oracle.default_params(2, noisy.shape), "auto-fast")
    _assert_identical(stats)


def test_volume_48_matches_oracle():
    print("48^3 volume", flush=True)
    _, noisy = phantoms.ellipsoid_volume((48, 48, 48), seed=52, peak=10.0, n_ellipsoids=12,
                                         axes_range=(3.0, 16.0))
    prof = BM4DProfile(noise_model=NoiseModel.POISSON)
    stats, _, _ = _compare(noisy, prof, oracle.default_params(3, noisy.shape), "auto-fast")
    _assert_identical(stats)


def test_128_volume_deterministic_and_batch_of_one():
    print("128^3 volume: determinism and batch of one", flush=True)
    _, noisy = phantoms.ellipsoid_volume((128, 128, 128), seed=53, peak=10.0, n_ellipsoids=20,
                                         axes_range=(4.0, 30.0))
    dev = torch.device("cuda", 0)
    prof = BM4DProfile(noise_model=NoiseModel.POISSON)
    pipe = PoissonPipeline(prof, noisy.shape, dev)
    v = torch.from_numpy(noisy).to(dev)
    ht1, wi1 = pipe.run(v)
    ht2, wi2 = pipe.run(v)
    assert torch.equal(wi1.est.view(torch.int32), wi2.est.view(torch.int32))
    assert torch.equal(wi1.sel, wi2.sel) and torch.equal(ht1.sel, ht2.sel)
    bp = BatchPipeline(prof, [noisy.shape], dev)
    bht, bwi = bp.run(bp.pack([noisy]))
    assert torch.equal(bwi.est.view(torch.int32), wi1.est.view(-1).view(torch.int32))
    assert torch.equal(bwi.sel, wi1.sel) and torch.equal(bht.sel, ht1.sel)
    assert torch.equal(bwi.counts, wi1.counts)
