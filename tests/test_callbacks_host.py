"""Callback order of bmnd() without a GPU: a stand-in pipeline records which
stage ran, so the tests show where each callback fires relative to the
stages.  The reference calls blockmatch_callback("ht") after stage-1
matching (bmndalgo.py:325-326), stage_callback("ht") after stage 1
(:632-633), blockmatch_callback("wiener") after stage-2 matching (:742-743)
and stage_callback("wiener") at the end (:1045-1046); a callback that raises
therefore stops the stages after it."""

from dataclasses import dataclass

import numpy as np
import pytest
import torch

import bmnd_amd.bmndalgo as bmndalgo
from bmnd_amd import BM3DProfile, NoiseModel
from bmnd_amd.shifts import get_reference_schedule_arrays


@dataclass
class _Out:
    est: torch.Tensor


class _Plan:
    def __init__(self, shape, block):
        self.block = block
        counts = tuple(shape[d] - block[d] + 1 for d in range(len(shape)))
        self.schedule = get_reference_schedule_arrays(counts, block, (3, 3))


class _FakePipeline:
    """Runs no kernels: each stage reports itself in `log` and calls its
    matching hook with one trivial group per reference."""

    def __init__(self, shape, log):
        self.plans = {"ht": _Plan(shape, (8, 8)), "wiener": _Plan(shape, (8, 8))}
        self.log = log

    def _match(self, stage, on_match):
        n = len(self.plans[stage].schedule.reference_abs)
        if on_match is not None:
            sel = torch.zeros((n, 16), dtype=torch.int32) - 1
            sel[:, 0] = 0
            on_match(sel, torch.ones(n, dtype=torch.int32), torch.ones(n, dtype=torch.int32))

    def run_ht(self, vol, timer=None, sigma_maps=False, on_match=None):
        self.log.append("ht.match")
        self._match("ht", on_match)
        self.log.append("ht.filter")
        return _Out(vol + 1.0), vol.clone(), None

    def run_wiener(self, vol, pilot, var1, timer=None, global_sigma=None, on_match=None):
        self.log.append("wiener.match")
        self._match("wiener", on_match)
        self.log.append("wiener.filter")
        return _Out(pilot + 1.0)


@pytest.fixture
def fake(monkeypatch):
    log = []
    monkeypatch.setattr(bmndalgo, "_device", lambda: torch.device("cpu"))
    monkeypatch.setattr(bmndalgo, "_pipeline",
                        lambda profile, shape, device: _FakePipeline(shape, log))
    return log


def _image():
    return np.random.default_rng(0).poisson(5.0, (32, 32)).astype(np.float32)


def test_callbacks_fire_between_the_stages(fake):
    seen = []
    out = bmndalgo.bmnd(_image(), BM3DProfile(noise_model=NoiseModel.POISSON),
                        stage_callback=lambda s, a: seen.append(("stage", s, float(a.mean()))),
                        blockmatch_callback=lambda s, r: seen.append(("match", s, len(r.groups))))
    order = [e[:2] for e in seen]
    assert order == [("match", "ht"), ("stage", "ht"), ("match", "wiener"), ("stage", "wiener")]
    # each callback sees the log state at its own point
    assert fake == ["ht.match", "ht.filter", "wiener.match", "wiener.filter"]
    assert seen[1][2] == pytest.approx(float(_image().mean()) + 1.0)
    assert out.dtype == np.float32


class _Stop(Exception):
    pass


def test_raising_ht_stage_callback_stops_stage_two(fake):
    def stop(stage, arr):
        if stage == "ht":
            raise _Stop
    with pytest.raises(_Stop):
        bmndalgo.bmnd(_image(), BM3DProfile(noise_model=NoiseModel.POISSON), stage_callback=stop)
    assert fake == ["ht.match", "ht.filter"]


def test_raising_ht_blockmatch_callback_stops_filtering(fake):
    def stop(stage, result):
        raise _Stop
    with pytest.raises(_Stop):
        bmndalgo.bmnd(_image(), BM3DProfile(noise_model=NoiseModel.POISSON),
                      blockmatch_callback=stop)
    assert fake == ["ht.match"]
