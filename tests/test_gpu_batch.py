"""Batches of inputs of uneven shapes (bmnd_batch / engine.BatchPipeline):
every kernel runs once over all inputs' references or tiles, and every
input's result must equal the single-input bmnd() result bit for bit, with
identical match lists in both stages; the small inputs are also checked
against the CPU oracle."""

import numpy as np
import pytest
import torch

import oracle
from bmnd_amd import BM3DProfile, BM4DProfile, NoiseModel, bmnd, bmnd_batch, phantoms
from bmnd_amd.engine import BatchPipeline, PoissonPipeline

pytestmark = pytest.mark.gpu


def _images():
    rng = np.random.default_rng(41)
    out = [phantoms.square_image(0)[1]]
    for side in ((37, 53), (96, 120), (8, 8), (40, 100), (71, 64)):
        _, img = phantoms.blob_image(max(max(side), 32), seed=int(rng.integers(1000)),
                                     peak=float(rng.choice([1.0, 5.0, 20.0, 100.0])))
        out.append(np.ascontiguousarray(img[:side[0], :side[1]]))
    return out


def _volumes():
    out = []
    for shape, seed in (((20, 24, 28), 1), ((24, 26, 22), 2), ((7, 36, 33), 3), ((16, 16, 16), 4)):
        _, v = phantoms.ellipsoid_volume(shape, seed=seed, peak=10.0, n_ellipsoids=5,
                                         axes_range=(2.0, 10.0))
        out.append(v)
    return out


def _same_bits(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.int32), b.view(np.int32))


def _check_batch_against_singles(inputs, prof):
    print(f"batch of {len(inputs)}: {[x.shape for x in inputs]}", flush=True)
    dev = torch.device("cuda", 0)
    bp = BatchPipeline(prof, [x.shape for x in inputs], dev)
    ht, wi = bp.run(bp.pack(inputs))
    outs = bp.unpack(wi.est)
    pilots = bp.unpack(ht.est)
    for i, x in enumerate(inputs):
        single = PoissonPipeline(prof, x.shape, dev)
        s_ht, s_wi = single.run(torch.from_numpy(np.ascontiguousarray(x)).to(dev))
        assert _same_bits(outs[i], s_wi.est.cpu().numpy()), f"input {i} output"
        assert _same_bits(pilots[i], s_ht.est.cpu().numpy()), f"input {i} pilot"
        for st_b, st_s in ((ht, s_ht), (wi, s_wi)):
            r0, r1 = st_b.stage.ref_off[i], st_b.stage.ref_off[i + 1]
            assert torch.equal(st_b.counts[r0:r1], st_s.counts), f"input {i} counts"
            assert torch.equal(st_b.sel[r0:r1], st_s.sel), f"input {i} lists"
    return outs


def test_batch_of_uneven_images_matches_single_calls():
    _check_batch_against_singles(_images(), BM3DProfile(noise_model=NoiseModel.POISSON))


def test_batch_of_uneven_volumes_matches_single_calls():
    _check_batch_against_singles(_volumes(), BM4DProfile(noise_model=NoiseModel.POISSON))


def test_batch_api_against_oracle_and_bmnd():
    """bmnd_batch on images (one all-zero, one with non-integer counts so the
    batch matches with the float pass) against the oracle and bmnd()."""
    imgs = _images()[:4]
    imgs.insert(2, np.zeros((30, 31), np.float32))
    imgs[1] = (imgs[1] * np.float32(0.37)).astype(np.float32)
    prof = BM3DProfile(noise_model=NoiseModel.POISSON)
    print("bmnd_batch: 5 images", flush=True)
    outs = bmnd_batch(imgs, prof)
    assert len(outs) == len(imgs)
    for i, x in enumerate(imgs):
        assert _same_bits(outs[i], bmnd(x, prof)), f"input {i} vs bmnd()"
        ref, _ = oracle.bmnd_poisson(x, oracle.default_params(2, x.shape))
        assert _same_bits(outs[i], ref), f"input {i} vs oracle"

