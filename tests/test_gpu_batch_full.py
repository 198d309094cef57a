"""The batched workloads at their real sizes (the inputs are the uneven
batch of phantoms.uneven_batch(0): 32 images with sides in [96, 1024] and 16
volumes with sides in [48, 128], peaks in [1, 50]; and sixteen 512^2 blob
images at peaks {1, 5, 20, 100}):
- each batch run twice gives the same estimates bit for bit;
- every input's match lists keep the structural invariants of the
  reference's finalisation (blockmatching.py:913-954: reference first,
  power-of-two sizes, distinct members inside the search box), checked per
  input of the flat list;
- the smallest and the largest input of each dimensionality equal the
  single-input pipeline on that input alone, lists and estimates;
- the smallest input of each dimensionality (a 281x141 image and a 73x49x55
  volume: what the CPU oracle finishes in seconds) equals the oracle bit for
  bit.
The per-input schedules and patch-grid indexing follow
bmndalgo.py:241-248, 676-686 per input."""

from types import SimpleNamespace

import numpy as np
import pytest
import torch

import oracle
from bmnd_amd import BM3DProfile, BM4DProfile, NoiseModel, bmnd_batch, phantoms
from bmnd_amd.engine import BatchPipeline, PoissonPipeline, table_first_order
from test_gpu_parity import _check_groups

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)


def _same_bits(a, b):
    return a.shape == b.shape and np.array_equal(np.asarray(a).view(np.int32),
                                                 np.asarray(b).view(np.int32))


def _run_batch(prof, inputs):
    order = table_first_order(inputs)
    inputs = [inputs[i] for i in order]
    bp = BatchPipeline(prof, [x.shape for x in inputs], DEV)
    packed = bp.pack(inputs)
    ht, wi = bp.run(packed)
    return bp, inputs, ht, wi


def _per_input(bp, out, name, i):
    st = bp.stages[name]
    r0, r1 = int(st.ref_off[i]), int(st.ref_off[i + 1])
    return SimpleNamespace(sel=out.sel[r0:r1], counts=out.counts[r0:r1],
                           accepted=out.accepted[r0:r1], schedule=st.schedules[i])


def _check(prof, inputs, with_oracle):
    bp, inputs, ht, wi = _run_batch(prof, inputs)
    est = [x.copy() for x in bp.unpack(wi.est)]
    pil = [x.copy() for x in bp.unpack(ht.est)]
    # determinism
    _, wi2 = bp.run(bp.pack(inputs))
    assert all(_same_bits(a, b) for a, b in zip(est, bp.unpack(wi2.est)))
    # invariants of every input's lists
    for i, x in enumerate(inputs):
        for name, out in (("ht", ht), ("wiener", wi)):
            plan = bp.stages[name].plan
            _check_groups(_per_input(bp, out, name, i), plan, x.shape,
                          plan.radius3[3 - x.ndim:])
    # the extreme inputs against the single-input pipeline
    sizes = [x.size for x in inputs]
    for i in sorted({int(np.argmin(sizes)), int(np.argmax(sizes))}):
        x = inputs[i]
        print(f"  single-input check of {x.shape}", flush=True)
        pipe = PoissonPipeline(prof, x.shape, DEV)
        s_ht, s_wi = pipe.run(torch.from_numpy(np.ascontiguousarray(x)).to(DEV))
        assert _same_bits(pil[i], s_ht.est.cpu().numpy())
        assert _same_bits(est[i], s_wi.est.cpu().numpy())
        for name, s_out, b_out in (("ht", s_ht, ht), ("wiener", s_wi, wi)):
            b = _per_input(bp, b_out, name, i)
            assert torch.equal(b.sel, s_out.sel) and torch.equal(b.counts, s_out.counts)
            assert torch.equal(b.accepted, s_out.accepted)
    # the smallest input against the oracle
    if with_oracle:
        i = int(np.argmin(sizes))
        print(f"  oracle check of {inputs[i].shape}", flush=True)
        o_out, o_pilot = oracle.bmnd_poisson(inputs[i], nthreads=16)
        assert _same_bits(pil[i], o_pilot) and _same_bits(est[i], o_out)


def test_uneven_batch_at_its_real_size():
    imgs, _, vols, _ = phantoms.uneven_batch(0)
    assert len(imgs) == 32 and len(vols) == 16
    assert all(96 <= s <= 1024 for x in imgs for s in x.shape)
    assert all(48 <= s <= 128 for x in vols for s in x.shape)
    print(f"uneven batch: 32 images {min(x.size for x in imgs)}..{max(x.size for x in imgs)} "
          f"voxels, 16 volumes {min(x.size for x in vols)}..{max(x.size for x in vols)}",
          flush=True)
    _check(BM3DProfile(noise_model=NoiseModel.POISSON), imgs, True)
    _check(BM4DProfile(noise_model=NoiseModel.POISSON), vols, True)


def test_sixteen_512_images_as_one_batch():
    pairs = [phantoms.blob_image(512, seed=s) for s in range(16)]
    noisy = [p[1] for p in pairs]
    print("16 x 512^2 batch", flush=True)
    _check(BM3DProfile(noise_model=NoiseModel.POISSON), noisy, False)
    # and through the public entry point
    outs = bmnd_batch(noisy[:3], BM3DProfile(noise_model=NoiseModel.POISSON))
    pipe = PoissonPipeline(BM3DProfile(noise_model=NoiseModel.POISSON), noisy[0].shape, DEV)
    for x, o in zip(noisy[:3], outs):
        _, wi = pipe.run(torch.from_numpy(x).to(DEV))
        assert _same_bits(o, wi.est.cpu().numpy())
