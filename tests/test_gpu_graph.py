"""Both stages captured as one HIP graph (bmnd_hip_pipeline_*,
engine.StageGraph) against the call-by-call pipeline: the same output and
match lists bit for bit on several different inputs launched one after
another with the program's own device work and synchronisations between the
launches; a new capture when an input breaks the count-table precondition
the graph was captured with (larger counts, non-integer values); non-default
options; and bmnd() through the graph equal to bmnd() call by call."""

import numpy as np
import pytest
import torch

import bmnd_amd.bmndalgo as bmndalgo
from bmnd_amd import BM3DProfile, BM4DProfile, NoiseModel, bmnd, phantoms
from bmnd_amd.engine import PoissonPipeline

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)


def _same(a, b):
    return torch.equal(a.contiguous().view(torch.int32), b.contiguous().view(torch.int32))


def _inputs(kind):
    if kind == "image":
        return BM3DProfile(noise_model=NoiseModel.POISSON), [
            phantoms.square_image(s)[1] for s in (0, 1, 2)]
    return BM4DProfile(noise_model=NoiseModel.POISSON), [
        phantoms.ellipsoid_volume((40, 44, 36), seed=s, peak=8.0, n_ellipsoids=8,
                                  axes_range=(3.0, 12.0))[1] for s in (81, 82, 83)]


@pytest.mark.parametrize("kind", ["image", "volume"])
def test_graph_launches_equal_the_call_by_call_pipeline(kind):
    print(f"graph vs eager: {kind}", flush=True)
    prof, inputs = _inputs(kind)
    pipe = PoissonPipeline(prof, inputs[0].shape, DEV)
    for rep in range(2):
        for x in inputs:
            v = torch.from_numpy(np.ascontiguousarray(x)).to(DEV)
            out = pipe.run_graph(v).clone()
            g = pipe._graph
            # the program's own work between launches
            ht, wi = pipe.run(v)
            torch.cuda.synchronize()
            assert _same(out, wi.est)
            assert _same(g.pilot, ht.est)
            sel, cnt, acc = g.lists["wiener"]
            assert torch.equal(sel.view(wi.sel.shape), wi.sel) and torch.equal(cnt, wi.counts)
            sel, cnt, acc = g.lists["ht"]
            assert torch.equal(sel.view(ht.sel.shape), ht.sel) and torch.equal(acc, ht.accepted)
    assert pipe._graph.modes[0] == 2  # integer counts: the count table in stage 1


def test_graph_recaptured_when_the_count_table_no_longer_fits():
    print("graph recapture", flush=True)
    prof = BM4DProfile(noise_model=NoiseModel.POISSON)
    low = phantoms.ellipsoid_volume((32, 36, 28), seed=91, peak=3.0, n_ellipsoids=6,
                                    axes_range=(3.0, 10.0))[1]
    high = phantoms.ellipsoid_volume((32, 36, 28), seed=92, peak=40.0, n_ellipsoids=6,
                                     axes_range=(3.0, 10.0))[1]
    frac = (low + 0.25).astype(np.float32)
    pipe = PoissonPipeline(prof, low.shape, DEV)
    eager = PoissonPipeline(prof, low.shape, DEV)
    # (input, stage-1 mode of the graph after the launch, its table size):
    # larger counts need a larger table, fractions the fast matcher, which
    # also serves integer counts (no new capture)
    for x, mode, tmax in ((low, 2, int(low.max())), (high, 2, int(high.max())),
                          (frac, 4, 0), (low, 4, 0)):
        v = torch.from_numpy(np.ascontiguousarray(x)).to(DEV)
        out = pipe.run_graph(v).clone()
        _, wi = eager.run(v)
        assert _same(out, wi.est)
        assert pipe._graph.modes[:2] == (mode, tmax)


@pytest.mark.parametrize("opts", [{"poisson_wiener_match_source": "observation"},
                                  {"poisson_variance_source_wiener": "observation"},
                                  {"poisson_wiener_gain_mode": "classic"},
                                  {"blockmatch_distance": "pearson"}])
def test_graph_with_options(opts):
    print(f"graph options {opts}", flush=True)
    prof = BM3DProfile(noise_model=NoiseModel.POISSON)
    for k, v in opts.items():
        setattr(prof, k, v)
    x = phantoms.blob_image(72, seed=93, peak=12.0)[1]
    pipe = PoissonPipeline(prof, x.shape, DEV)
    v = torch.from_numpy(np.ascontiguousarray(x)).to(DEV)
    out = pipe.run_graph(v).clone()
    _, wi = pipe.run(v)
    assert _same(out, wi.est)


def test_bmnd_through_the_graph_equals_call_by_call(monkeypatch):
    print("bmnd graph", flush=True)
    x = phantoms.square_image(5)[1]
    prof = BM3DProfile(noise_model=NoiseModel.POISSON)
    monkeypatch.setattr(bmndalgo, "USE_GRAPHS", True)
    a = bmnd(x, prof)
    b = bmnd(x, prof)
    monkeypatch.setattr(bmndalgo, "USE_GRAPHS", False)
    c = bmnd(x, BM3DProfile(noise_model=NoiseModel.POISSON))
    assert np.array_equal(a.view(np.int32), c.view(np.int32))
    assert np.array_equal(b.view(np.int32), c.view(np.int32))


@pytest.mark.parametrize("kind", ["image", "volume"])
def test_every_aggregation_tile_gives_the_same_estimate(kind):
    """The aggregation tiles the tuner chooses among (512, 256 and 128
    voxels) sum each voxel in the same order: identical estimates."""
    print(f"aggregation tiles: {kind}", flush=True)
    from bmnd_amd import _lib
    prof, inputs = _inputs(kind)
    pipe = PoissonPipeline(prof, inputs[0].shape, DEV)
    v = torch.from_numpy(np.ascontiguousarray(inputs[0])).to(DEV)
    L = _lib.lib()
    outs = []
    try:
        for t in (0, 1, 2):
            _lib.check(L.bmnd_hip_aggregate_force_tile(t))
            ht, wi = pipe.run(v)
            outs.append((ht.est.clone(), wi.est.clone()))
    finally:
        _lib.check(L.bmnd_hip_aggregate_force_tile(-1))
    ht, wi = pipe.run(v)  # tuned
    for a, b in outs:
        assert _same(a, ht.est) and _same(b, wi.est)
