"""Pins of the oracle's group-mass conservation (poisson.py:49-265) on the
CPU, from the reference's own tests and from what the construction fixes:
- one group's correction (tests/test_poisson_unit.py:28-112): with the DCT
  and bior1.5 synthesis of an 8x8 group, the correction changes the
  synthesised group by a constant, the corrected group carries the target
  sum(O * source), and scaling the functional or the direction changes
  nothing; the per-axis directions synthesise ones;
- the pipeline (tests/test_bmnd_pipeline_unit.py:668-702): with the HT,
  Wiener or both stages conserving, the corrected stage's estimate keeps the
  observation's total (the normalised occurrence weights of every voxel sum
  to one, so the per-group targets add up to the total), on the reference's
  8x8 Poisson(3) image, a 26x24 image and a small volume, where the plain
  stages do not keep it; with a stage not conserving,
  that stage is the plain one bit for bit."""

import numpy as np
import pytest

import oracle


def _matrices(kind):
    if kind == "dct":
        f, i = oracle.dct(8)
    else:
        f, i = oracle.bior15(8)
        i = np.asarray(np.linalg.inv(f), dtype=np.float32)
    return f, i


def _synth(x, mats):
    return oracle.apply_transform(np.ascontiguousarray(x, np.float32), mats)


def _group(kind):
    f, i = _matrices(kind)
    source = np.arange(1, 65, dtype=np.float32).reshape(8, 8)
    O = np.linspace(0.1, 0.8, 64).reshape(8, 8)
    filtered = (np.float32(0.5) * oracle.apply_transform(source, [f, f])).astype(np.float32)
    dirs = [np.linalg.solve(i.astype(np.float64), np.ones(8))] * 2
    return source, O, filtered, [i, i], dirs


@pytest.mark.parametrize("kind", ["dct", "bior"])
def test_correction_is_constant_and_exact(kind):
    source, O, filtered, inv, dirs = _group(kind)
    corrected = oracle.mass_correct(filtered, source, O, inv, dirs)
    synth = _synth(corrected, inv)
    delta = synth - _synth(filtered, inv)
    assert corrected.dtype == np.float32
    assert np.allclose(delta, delta.reshape(-1)[0], atol=1e-5)
    assert np.isclose(np.sum(O * synth), np.sum(O * source), rtol=1e-6)


@pytest.mark.parametrize("kind", ["dct", "bior"])
@pytest.mark.parametrize("fscale,dscale", [(1e-12, 1.0), (1.0, 1e-9), (1e9, 1e6)])
def test_correction_is_invariant_to_scaling(kind, fscale, dscale):
    source, O, filtered, inv, dirs = _group(kind)
    a = oracle.mass_correct(filtered, source, O, inv, dirs)
    b = oracle.mass_correct(filtered, source, O * fscale, inv, [dirs[0] * dscale, dirs[1]])
    assert np.allclose(b, a, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("kind", ["dct", "bior"])
def test_constant_direction_synthesises_ones(kind):
    _, i = _matrices(kind)
    d = np.linalg.solve(i.astype(np.float64), np.ones(8))
    assert np.allclose(i.astype(np.float64) @ d, 1.0, atol=1e-6)


def _run(vol, mode, nthreads=4):
    p = oracle.default_params(vol.ndim, vol.shape, nthreads=nthreads)
    oracle.set_mass_conservation(p, mode)
    return oracle.bmnd_poisson(vol, p)


@pytest.mark.parametrize("mode", ["ht", "wiener", "both"])
@pytest.mark.parametrize("kind", ["reference-image", "image", "volume"])
def test_conserving_stages_keep_the_total(kind, mode):
    rng = np.random.default_rng(43)
    if kind == "reference-image":
        vol = rng.poisson(3.0, size=(8, 8)).astype(np.float32)
    elif kind == "image":
        truth = np.full((26, 24), 3.0)
        truth[6:18, 5:15] = 9.0
        vol = rng.poisson(truth).astype(np.float32)
    else:
        truth = np.full((14, 15, 13), 3.0)
        truth[3:10, 4:11, 2:9] = 12.0
        vol = rng.poisson(truth).astype(np.float32)
    out, pilot = _run(vol, mode)
    plain_out, plain_pilot = _run(vol, "none")
    total = np.sum(vol, dtype=np.float64)
    if mode in ("ht", "both"):
        assert np.isclose(np.sum(pilot, dtype=np.float64), total, rtol=2e-6)
        if kind != "reference-image":  # one 8x8 patch keeps its total anyway
            assert not np.isclose(np.sum(plain_pilot, dtype=np.float64), total, rtol=2e-6)
    else:
        assert np.array_equal(pilot.view(np.int32), plain_pilot.view(np.int32))
    if mode in ("wiener", "both"):
        assert np.isclose(np.sum(out, dtype=np.float64), total, rtol=2e-6)
        # the plain stage does not keep it (the check above is not vacuous)
        if kind != "reference-image":
            assert not np.isclose(np.sum(plain_out, dtype=np.float64), total, rtol=2e-6)
