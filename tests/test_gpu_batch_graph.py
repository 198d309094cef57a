"""Several batches captured as one HIP graph (bmnd_hip_pipeline_create_many,
engine.BatchGraph) against the call-by-call BatchPipeline: a batch of
uneven images and a batch of uneven volumes in one graph, launched on
different inputs one after another with the program's own device work and
synchronisations between the launches, give the same estimates, pilots and
match lists bit for bit; an input that breaks the count-table precondition
the graph was captured with makes a new capture."""

import numpy as np
import pytest
import torch

from bmnd_amd import BM3DProfile, BM4DProfile, NoiseModel, phantoms
from bmnd_amd.engine import BatchGraph, BatchPipeline

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)


def _images(seed):
    rng = np.random.default_rng(seed)
    out = []
    for side in ((37, 53), (64, 80), (8, 8), (40, 100)):
        _, img = phantoms.blob_image(max(max(side), 32), seed=int(rng.integers(1000)),
                                     peak=float(rng.choice([1.0, 5.0, 20.0])))
        out.append(np.ascontiguousarray(img[:side[0], :side[1]]))
    return out


def _volumes(seed):
    out = []
    for i, shape in enumerate(((20, 24, 28), (24, 26, 22), (7, 36, 33))):
        _, v = phantoms.ellipsoid_volume(shape, seed=seed + i, peak=8.0, n_ellipsoids=5,
                                         axes_range=(2.0, 10.0))
        out.append(v)
    return out


def _same(a, b):
    return torch.equal(a.contiguous().view(torch.int32), b.contiguous().view(torch.int32))


def _check(g, j, bp, packed):
    ht, wi = bp.run(packed)
    torch.cuda.synchronize()
    assert _same(g.outputs[j], wi.est), f"batch {j} output"
    assert _same(g.pilots[j], ht.est), f"batch {j} pilot"
    for name, out in (("ht", ht), ("wiener", wi)):
        sel, cnt, acc = g.lists[j][name]
        assert torch.equal(sel.view(out.sel.shape), out.sel), f"batch {j} {name} lists"
        assert torch.equal(cnt, out.counts) and torch.equal(acc, out.accepted)


def test_images_and_volumes_in_one_graph_equal_call_by_call():
    i2, i3 = _images(51), _volumes(61)
    bp2 = BatchPipeline(BM3DProfile(noise_model=NoiseModel.POISSON), [x.shape for x in i2], DEV)
    bp3 = BatchPipeline(BM4DProfile(noise_model=NoiseModel.POISSON), [x.shape for x in i3], DEV)
    g = BatchGraph([bp2, bp3])
    for rep, seeds in enumerate(((51, 61), (52, 62), (51, 61))):
        print(f"batch graph launch {rep}: seeds {seeds}", flush=True)
        imgs, vols = _images(seeds[0]), _volumes(seeds[1])
        p2, p3 = bp2.pack(imgs), bp3.pack(vols)
        g.run([p2, p3])
        outs = [o.clone() for o in g.outputs]
        _check(g, 0, bp2, p2)
        _check(g, 1, bp3, p3)
        assert all(_same(a, b) for a, b in zip(outs, g.outputs))
    assert [m[0] for m in g.modes] == [2, 2]  # integer counts: the count table in stage 1


def test_batch_graph_recaptured_when_the_count_table_no_longer_fits():
    prof = BM3DProfile(noise_model=NoiseModel.POISSON)
    low = _images(71)
    high = [x.copy() for x in low]
    high[1] = np.rint(high[1] * 9.0 + 120.0).astype(np.float32)  # counts above 100
    frac = [x.copy() for x in low]
    frac[2] = (frac[2] + 0.25).astype(np.float32)
    bp = BatchPipeline(prof, [x.shape for x in low], DEV)
    g = BatchGraph([bp])
    # (inputs, stage-1 mode of the graph after the launch): counts above the
    # table and fractions make the fast matcher, which also serves the
    # integer counts that follow (no new capture)
    for inputs, mode in ((low, 2), (high, 4), (low, 4), (frac, 4)):
        print(f"batch graph recapture: mode {mode}", flush=True)
        packed = bp.pack(inputs)
        g.run([packed])
        assert g.modes[0][0] == mode
        _check(g, 0, bp, packed)
